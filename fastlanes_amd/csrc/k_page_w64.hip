// page kernels for 64-bit elements: integer and patched-FFOR pages
#define FLS_W 64
#define FLS_PART 1
#include "k_page_w.inc"
