// page kernels for 8-bit elements: patched FFOR and DELTA pages
#define FLS_W 8
#define FLS_PART 3
#include "k_page_w.inc"
