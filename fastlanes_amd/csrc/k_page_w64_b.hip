// page kernels for 64-bit elements: RLE and dictionary pages (patched dictionary pages included)
#define FLS_W 64
#define FLS_PART 2
#include "k_page_w.inc"
