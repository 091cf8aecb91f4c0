// Instantiations of the batched search kernel: 256-thread workgroups, mixed-precision (FP32 pair arithmetic, FP64 accumulation)
// form, both lane counts and both species modes.  Each (workgroup size, precision) pair has its own
// translation unit so the variants compile in parallel.
#include "search_kernel.hpp"

namespace flyg {
  namespace sk {
    namespace {
      template <int L>
      void const* pick(bool multi) {
        return multi ? reinterpret_cast<void const*>(&k_search<256, L, true, true>)
                     : reinterpret_cast<void const*>(&k_search<256, L, true, false>);
      }
    }  // namespace

    void const* search_kernel_256_mixed(int L, bool multi) {
      if (L == 8) return pick<8>(multi);
      if (L == 16) return pick<16>(multi);
      return nullptr;
    }
  }  // namespace sk
}  // namespace flyg
