// C-linkage shim over the reference's own mmio.cpp (compiled unmodified from
// the reference tree into oracle/_ref/). Used only by tests to pin the
// oracle's Matrix Market banner and size parsing (mmio.cpp:96-217).
#include <stdio.h>
#include "mmio.h"

extern "C" int ref_mm_banner(const char *path, char *tc4)
{
    FILE *f = fopen(path, "r");
    if (!f) return -1;
    MM_typecode tc;
    int rc = mm_read_banner(f, &tc);
    for (int i = 0; i < 4; i++) tc4[i] = tc[i];
    fclose(f);
    return rc;
}

// Only for files whose size line is "M N nz": on other files the reference's
// fscanf loop (mmio.cpp:210-214) may never reach EOF.
extern "C" int ref_mm_banner_size(const char *path, char *tc4, int *M, int *N,
                                  unsigned long *nz, int *rc_size)
{
    FILE *f = fopen(path, "r");
    if (!f) return -1;
    MM_typecode tc;
    int rc = mm_read_banner(f, &tc);
    for (int i = 0; i < 4; i++) tc4[i] = tc[i];
    *rc_size = -1;
    if (rc == 0) *rc_size = mm_read_mtx_crd_size(f, M, N, nz);
    fclose(f);
    return rc;
}
