// page kernels for 64-bit elements: patched FFOR and DELTA pages
#define FLS_W 64
#define FLS_PART 3
#include "k_page_w.inc"
