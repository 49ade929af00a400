// page kernels for 8-bit elements: RLE and dictionary pages (patched dictionary pages included)
#define FLS_W 8
#define FLS_PART 2
#include "k_page_w.inc"
