// Device scan and stable radix sort shared by COO -> CSR and the Schur
// pattern / normal-equation topology builders.
#pragma once

#include "pm_common.hpp"

namespace pm
{

// Exclusive scan of n u32 values (in may equal out); `grand_total` (device,
// may be null) receives the sum.  `scratch` holds scan_scratch_words(n) words.
int64_t scan_scratch_words(int64_t n);
int scan_u32(const unsigned* in, unsigned* out, int64_t n, unsigned* scratch, unsigned* grand_total,
             hipStream_t stream);

// Stable LSD radix sort of (key, value) pairs on the low key_bits bits, 8 bits
// per pass.  The input is in keys_a / vals_a; the sorted result ends in the
// b buffers when *in_b is set, else in the a buffers.  `hist` holds
// radix_hist_words(count) words and `scratch` scan_scratch_words of that.
int radix_sort_pairs(unsigned long long* keys_a, unsigned* vals_a, unsigned long long* keys_b, unsigned* vals_b,
                     int64_t count, int key_bits, unsigned* hist, unsigned* scratch, hipStream_t stream, bool* in_b);
int64_t radix_hist_words(int64_t count);

// Smallest b with 2^b >= n.
int bits_for(int64_t n);

} // namespace pm
