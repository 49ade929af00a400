// page kernels for 64-bit elements: UNCOMPRESSED, FFOR and DELTA pages
#define FLS_W 64
#define FLS_PART 1
#include "k_page_w.inc"
