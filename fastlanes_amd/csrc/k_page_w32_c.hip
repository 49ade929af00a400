// page kernels for 32-bit elements: patched FFOR and DELTA pages
#define FLS_W 32
#define FLS_PART 3
#include "k_page_w.inc"
