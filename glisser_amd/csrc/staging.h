// Library-internal access to a context's pinned staging buffers, for the
// C++ host loop in the same library (not part of the C ABI).
#pragma once

#include <cstdint>

#include "glisser_amd.h"

namespace glx {

// glx_pack_encounters without the final copy: packs particles
// [begin, begin + n) of the context into its pinned staging buffer and sets
// *out to it.  The records stay valid until the next pack on this context.
int pack_encounters_staged(glx_ctx* c, uint64_t begin, uint64_t n, double t_block, const glx_encounter** out);

// The same in two halves: _begin packs the records on the context's main
// stream (after the work already queued there, before whatever is queued
// next) and copies them to the pinned staging buffer on a copy stream of
// their own, without waiting; _wait waits for that copy and sets *out to
// the buffer.  Work queued on the main stream after _begin may run while
// the copy does.
int pack_encounters_begin(glx_ctx* c, uint64_t begin, uint64_t n, double t_block);
int pack_encounters_wait(glx_ctx* c, const glx_encounter** out);

}  // namespace glx
