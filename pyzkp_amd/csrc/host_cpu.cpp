// CPU feature checks for the host-side SIMD code.  Built with no vector code
// generation flags (it runs before anyone knows what the CPU has), and with
// the host compiler, since the checks are host builtins.
#include "g2_simd.hpp"

namespace zk {
namespace g2simd {

bool cpu_has_ifma() {
  __builtin_cpu_init();
  return __builtin_cpu_supports("avx512f") && __builtin_cpu_supports("avx512ifma");
}

}  // namespace g2simd
}  // namespace zk
