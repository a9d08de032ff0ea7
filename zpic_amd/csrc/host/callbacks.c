/*
 * callbacks.c - adapter for field callbacks written in other languages.
 *
 * The em2d field callbacks return a float3 by value (em2d/emf.h:37-38,
 * 59-60), which foreign-function layers such as Python's ctypes cannot
 * produce.  Pass zpic_field_callback as E_custom/B_custom and a
 * zpic_field_cb as the matching *_custom_data: the adapter calls cb->fn,
 * which writes the three components through a pointer.
 */
#include "zpic_gpu.h"

float3 zpic_field_callback(int ix, float dx, int iy, float dy, void *data)
{
    const zpic_field_cb *cb = (const zpic_field_cb *)data;
    float v[3] = {0.0f, 0.0f, 0.0f};
    cb->fn(ix, dx, iy, dy, v, cb->user);
    float3 r = {v[0], v[1], v[2]};
    return r;
}
