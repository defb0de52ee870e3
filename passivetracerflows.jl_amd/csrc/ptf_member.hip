// The compiled whole-member step kernels (ptf_member.hpp): square 2D members of 32^2,
// 64^2 and (FP32) 128^2, whose two half-spectrum planes fit the 160 KiB of LDS.
#include "ptf_member.hpp"

namespace ptf {

namespace {
template <typename T, int N, int RC, int RR>
const void* member_entry(int* threads) {
  *threads = member_threads<N, RC, RR>();
  return reinterpret_cast<const void*>(&member_step_kernel<T, N, RC, RR>);
}
}  // namespace

const void* member_kernel(bool dbl, int n, int* threads) {
  if (dbl) {
    switch (n) {
      case 32: return member_entry<double, 32, 4, 4>(threads);
      case 64: return member_entry<double, 64, 4, 4>(threads);
      default: return nullptr;
    }
  }
  switch (n) {
    case 32: return member_entry<float, 32, 4, 4>(threads);
    case 64: return member_entry<float, 64, 4, 4>(threads);
    case 128: return member_entry<float, 128, 16, 16>(threads);
    default: return nullptr;
  }
}

}  // namespace ptf
