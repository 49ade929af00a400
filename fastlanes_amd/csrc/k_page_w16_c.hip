// page kernels for 16-bit elements: patched FFOR and DELTA pages
#define FLS_W 16
#define FLS_PART 3
#include "k_page_w.inc"
