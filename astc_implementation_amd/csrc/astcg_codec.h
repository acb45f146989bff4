// Host bitstream helpers shared by the C ABI.
#pragma once
#include <cstdint>

#include "../../include/astcg.h"

namespace astcg {

void encode_ise_host(int quant, int count, const uint8_t* in, uint8_t* out, int bit_offset);
void decode_ise_host(int quant, int count, const uint8_t* in, int bit_offset, uint8_t* out);
int symbolic_to_physical_host(int bx, int by, const astcg_symbolic_block& sb, uint8_t out[16]);
int physical_to_symbolic_host(int bx, int by, const uint8_t in[16], astcg_symbolic_block* sb);
int decode_block_rgba8(int bx, int by, const uint8_t in[16], uint8_t* texels);

}  // namespace astcg
