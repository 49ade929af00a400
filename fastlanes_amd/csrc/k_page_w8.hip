// page kernels for 8-bit elements: UNCOMPRESSED, FFOR and DELTA pages
#define FLS_W 8
#define FLS_PART 1
#include "k_page_w.inc"
