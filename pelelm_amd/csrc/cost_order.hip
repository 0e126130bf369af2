// The sort behind the cost ordering (mi.order): a stable ascending radix sort of
// (cost key, cell) pairs with hipCUB, so that each 64-cell wave of workspace slots holds
// cells of similar integration cost.  Kept in its own translation unit: it is the only
// code that needs hipCUB.
#include <hipcub/hipcub.hpp>

#include "reactor_internal.h"

namespace pelelm {

size_t cost_sort_temp_bytes(long n) {
  size_t bytes = 0;
  hipcub::DoubleBuffer<unsigned int> k(nullptr, nullptr);
  hipcub::DoubleBuffer<int> v(nullptr, nullptr);
  if (hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, k, v, (int)n, 0, 32) != hipSuccess) return 0;
  return bytes;
}

hipError_t cost_sort(void* temp, size_t temp_bytes, unsigned int* keys, int* vals, long n, hipStream_t stream,
                     const int** sorted_vals) {
  hipcub::DoubleBuffer<unsigned int> k(keys, keys + n);
  hipcub::DoubleBuffer<int> v(vals, vals + n);
  const hipError_t e = hipcub::DeviceRadixSort::SortPairs(temp, temp_bytes, k, v, (int)n, 0, 32, stream);
  if (e != hipSuccess) return e;
  *sorted_vals = v.Current();
  return hipSuccess;
}

}  // namespace pelelm
