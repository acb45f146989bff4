/* Oracle colour endpoint packing/unpacking (test infrastructure only). */
#ifndef OR_COLOR_H
#define OR_COLOR_H
#include "or_internal.h"

int or_pack_color_endpoints(const float* color0, const float* color1, const float* rgbs, int format, int q,
                            uint32_t* packed);
void or_unpack_color_endpoints(int format, const uint32_t* v, int* o0, int* o1);
void or_rgba_delta_unpack(const int* in0, const int* in1, int* o0, int* o1);
void or_rgba_unpack(const int* in0, const int* in1, int* o0, int* o1);

#endif
