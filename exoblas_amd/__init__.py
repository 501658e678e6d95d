"""exoblas_amd: a BLAS for AMD Instinct MI355X (gfx950) behind the exo_*
C entry points of ExoBLAS.

    exoblas_amd.lib      ctypes binding of libexoblas_amd.so (every exo_*
                         symbol and its device-pointer _async form)
    exoblas_amd.cblas    CBLAS-style functions mirroring the reference's
                         C++ wrappers (argument handling, early returns)
    exoblas_amd.generator  the exported symbol table; writes the headers and
                         the C-ABI glue
    exoblas_amd.build    compiles the HIP sources for gfx950, in-tree
"""
