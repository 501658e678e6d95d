/* Value types of the exo_* C-ABI.
 *
 * Window structs: a pointer to element 0 plus int_fast32_t element strides,
 * passed by value.  Field names (.data, .strides) and layout are the ones the
 * reference's CBLAS-style wrappers build with designated initialisers, e.g.
 * test/level1/axpy/exo_axpy_wrapper.h:14-15 and
 * test/level2/gemv/exo_gemv_wrapper.h:21-24.  On x86-64 glibc int_fast32_t is
 * 8 bytes, so a 1-D window is 16 bytes and a 2-D window 24 bytes.
 * The "c" forms are the read-only windows.
 *
 * Enum arguments cross the ABI as int_fast32_t carrying the netlib CBLAS
 * values (src/common/cblas_enums.py:4-33, test/common/cblas.h:10-18).
 */
#ifndef EXOBLAS_AMD_EXO_TYPES_H
#define EXOBLAS_AMD_EXO_TYPES_H

#include <stdint.h>

#ifndef EXO_WIN_1F32
#define EXO_WIN_1F32
struct exo_win_1f32 {
  float *const data;
  const int_fast32_t strides[1];
};
#endif
#ifndef EXO_WIN_1F32C
#define EXO_WIN_1F32C
struct exo_win_1f32c {
  const float *const data;
  const int_fast32_t strides[1];
};
#endif
#ifndef EXO_WIN_1F64
#define EXO_WIN_1F64
struct exo_win_1f64 {
  double *const data;
  const int_fast32_t strides[1];
};
#endif
#ifndef EXO_WIN_1F64C
#define EXO_WIN_1F64C
struct exo_win_1f64c {
  const double *const data;
  const int_fast32_t strides[1];
};
#endif
#ifndef EXO_WIN_2F32
#define EXO_WIN_2F32
struct exo_win_2f32 {
  float *const data;
  const int_fast32_t strides[2];
};
#endif
#ifndef EXO_WIN_2F32C
#define EXO_WIN_2F32C
struct exo_win_2f32c {
  const float *const data;
  const int_fast32_t strides[2];
};
#endif
#ifndef EXO_WIN_2F64
#define EXO_WIN_2F64
struct exo_win_2f64 {
  double *const data;
  const int_fast32_t strides[2];
};
#endif
#ifndef EXO_WIN_2F64C
#define EXO_WIN_2F64C
struct exo_win_2f64c {
  const double *const data;
  const int_fast32_t strides[2];
};
#endif

/* netlib CBLAS enum values */
#define EXOBLAS_NO_TRANS 111
#define EXOBLAS_TRANS 112
#define EXOBLAS_CONJ_TRANS 113
#define EXOBLAS_UPPER 121
#define EXOBLAS_LOWER 122
#define EXOBLAS_NON_UNIT 131
#define EXOBLAS_UNIT 132
#define EXOBLAS_LEFT 141
#define EXOBLAS_RIGHT 142

#endif /* EXOBLAS_AMD_EXO_TYPES_H */
