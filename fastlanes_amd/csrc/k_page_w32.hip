// page kernels for 32-bit elements: UNCOMPRESSED, FFOR and DELTA pages
#define FLS_W 32
#define FLS_PART 1
#include "k_page_w.inc"
