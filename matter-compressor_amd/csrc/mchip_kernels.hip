// mchip_kernels.hip -- the single device translation unit: the decode,
// encode and hash kernels share the constant tables (DCT matrix, scan
// order, band starts) uploaded once by mchip_init.
#include "mchip_decode.hip"
#include "mchip_encode.hip"
#include "mchip_hash.hip"
#include "mchip_deblock.hip"
