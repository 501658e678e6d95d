# CMake package for exoblas_amd: imported targets with the names the
# reference's test harness links against (EXO_BLAS::exo_<kernel>,
# test/CMakeLists.txt:42,72), all backed by the one shared library.
#
#   list(APPEND CMAKE_PREFIX_PATH /path/to/repo/exoblas_amd/cmake)
#   find_package(EXO_BLAS CONFIG REQUIRED)
#   target_link_libraries(saxpy_correctness PRIVATE EXO_BLAS::exo_axpy)
get_filename_component(_exoblas_pkg "${CMAKE_CURRENT_LIST_DIR}/.." ABSOLUTE)
get_filename_component(_exoblas_root "${_exoblas_pkg}/.." ABSOLUTE)
set(_exoblas_lib "${_exoblas_pkg}/libexoblas_amd.so")
if(NOT EXISTS "${_exoblas_lib}")
  message(FATAL_ERROR "exoblas_amd is not built: run `python -m exoblas_amd.build` in ${_exoblas_root}")
endif()
if(NOT TARGET EXO_BLAS::exoblas_amd)
  add_library(EXO_BLAS::exoblas_amd SHARED IMPORTED)
  set_target_properties(EXO_BLAS::exoblas_amd PROPERTIES
    IMPORTED_LOCATION "${_exoblas_lib}"
    INTERFACE_INCLUDE_DIRECTORIES "${_exoblas_root}/include;/opt/rocm/include"
    INTERFACE_COMPILE_DEFINITIONS "__HIP_PLATFORM_AMD__")
  foreach(_k asum axpy dot dsdot copy swap rot rotm scal gemv gbmv ger symv sbmv syr syr2 trmv tbmv trsv tbsv
             gemm mscal symm syrk syr2k trmm)
    add_library(EXO_BLAS::exo_${_k} INTERFACE IMPORTED)
    set_target_properties(EXO_BLAS::exo_${_k} PROPERTIES INTERFACE_LINK_LIBRARIES EXO_BLAS::exoblas_amd)
  endforeach()
endif()
set(EXO_BLAS_FOUND TRUE)
