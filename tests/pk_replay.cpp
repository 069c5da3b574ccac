// Host replay of the packed 16-bit LDS tile's accounting (pk_count.h), built
// by tests/test_pk_count.py with a host C++ compiler.
//
// A schedule is a list of adds (bin, inc) in the order their atomics reach the
// word, plus for each add the step after which its lane accounts for it (any
// step at or after its own: lanes may be descheduled for arbitrarily long
// between the atomic and the accounting).  The words take the adds as 32-bit
// wrapping sums, the tallies take what pk_owed says each add owes, and the
// result is each bin's pk_read.
#include <stdint.h>

#include <algorithm>
#include <numeric>
#include <vector>

#include "pk_count.h"

extern "C" int pk_replay(const uint32_t* bin, const uint32_t* inc, const int64_t* account_at,
                         int64_t n, uint32_t width, const uint32_t* init_words,
                         uint64_t* counts_out) {
  const uint32_t nwords = (width + 1u) / 2u;
  std::vector<uint32_t> word(init_words, init_words + nwords);
  std::vector<uint32_t> tally(width + 1u, 0u);
  std::vector<uint32_t> returned(n);
  // accounting events ordered by the step after which they happen
  std::vector<int64_t> order(n);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(),
                   [&](int64_t a, int64_t b) { return account_at[a] < account_at[b]; });
  int64_t next = 0;
  for (int64_t k = 0; k < n; k++) {
    if (bin[k] >= width || inc[k] == 0 || inc[k] >= 0x10000u || account_at[k] < k) return -1;
    uint32_t& w = word[bin[k] >> 1];
    returned[k] = w;
    w += inc[k] << (16u * (bin[k] & 1u));
    while (next < n && account_at[order[next]] <= k) {
      const int64_t a = order[next++];
      if (a > k) return -2;  // accounted before its own atomic
      const bale::PkOwed o = bale::pk_owed(returned[a], bin[a] & 1u, inc[a]);
      // the kernels skip pk_owed unless pk_wrapped: it must owe nothing then
      if (!bale::pk_wrapped(returned[a], bin[a] & 1u, inc[a]) && (o.own || o.high)) return -4;
      tally[bin[a]] += o.own;
      if (o.high && (bin[a] | 1u) < width) tally[bin[a] | 1u] += o.high;
    }
  }
  if (next != n) return -3;
  for (uint32_t b = 0; b < width; b++)
    counts_out[b] = bale::pk_read(word[b >> 1], b & 1u, tally[b]);
  return 0;
}
