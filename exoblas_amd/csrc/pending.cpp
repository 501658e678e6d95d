// Entry points whose kernels are not built yet: they report kNotSupported.
#include "ops.hpp"

namespace exoblas {
namespace ops {

[[noreturn]] static void pending(const char* what) { throw Error(kNotSupported, std::string(what) + " is not implemented yet"); }

template <class T>
void trsv_rm(Ctx&, int64_t, int64_t, int64_t, int64_t, T*, int64_t, const T*, int64_t) { pending("trsv"); }
template <class T>
void gemm(Ctx&, int64_t, int64_t, int64_t, int64_t, int64_t, T, const T*, int64_t, const T*, int64_t, const T*,
          int64_t, const T*, int64_t, T*, int64_t) { pending("gemm"); }
template <class T> void mscal_rm(Ctx&, int64_t, int64_t, T, T*, int64_t) { pending("mscal"); }
template <class T> void trmscal_rm(Ctx&, int64_t, int64_t, T, T*, int64_t) { pending("trmscal"); }
template <class T>
void symm_rm(Ctx&, int64_t, int64_t, int64_t, int64_t, T, const T*, int64_t, const T*, int64_t, const T*, int64_t,
             T*, int64_t) { pending("symm"); }
template <class T>
void syrk_rm(Ctx&, int64_t, int64_t, int64_t, int64_t, T, const T*, int64_t, const T*, int64_t, const T*, int64_t,
             const T*, int64_t, T*, int64_t) { pending("syrk"); }

#define INST(T)                                                                                               \
  template void trsv_rm<T>(Ctx&, int64_t, int64_t, int64_t, int64_t, T*, int64_t, const T*, int64_t);         \
  template void gemm<T>(Ctx&, int64_t, int64_t, int64_t, int64_t, int64_t, T, const T*, int64_t, const T*,    \
                        int64_t, const T*, int64_t, const T*, int64_t, T*, int64_t);                          \
  template void mscal_rm<T>(Ctx&, int64_t, int64_t, T, T*, int64_t);                                          \
  template void trmscal_rm<T>(Ctx&, int64_t, int64_t, T, T*, int64_t);                                        \
  template void symm_rm<T>(Ctx&, int64_t, int64_t, int64_t, int64_t, T, const T*, int64_t, const T*, int64_t, \
                           const T*, int64_t, T*, int64_t);                                                   \
  template void syrk_rm<T>(Ctx&, int64_t, int64_t, int64_t, int64_t, T, const T*, int64_t, const T*, int64_t, \
                           const T*, int64_t, const T*, int64_t, T*, int64_t);
INST(float)
INST(double)

}  // namespace ops
}  // namespace exoblas
