// Host check of the staging path's copy pool: many sizes (including ones
// that leave some threads without a share), from two host threads at once;
// every byte of every destination must equal its source and nothing past
// the end may change.
#include <cstdio>
#include <cstring>
#include <random>
#include <thread>
#include <vector>

#include "copy_pool.hpp"

int main() {
  exoblas::CopyPool pool(4);
  int failures = 0;
  auto worker = [&](unsigned seed) {
    std::mt19937_64 rng(seed);
    for (int it = 0; it < 200; ++it) {
      const size_t bytes = (it % 7 == 0) ? (size_t)(rng() % 64) : (size_t)(rng() % (3 << 20));
      std::vector<unsigned char> src(bytes + 64), dst(bytes + 64, 0xA5);
      for (auto& b : src) b = (unsigned char)rng();
      pool.copy(dst.data(), src.data(), bytes);
      if (std::memcmp(dst.data(), src.data(), bytes) != 0) ++failures;
      for (size_t i = bytes; i < bytes + 64; ++i)
        if (dst[i] != 0xA5) {
          ++failures;
          break;
        }
    }
  };
  std::thread a(worker, 1u), b(worker, 2u);
  a.join();
  b.join();
  std::printf("%s %d\n", failures ? "FAIL" : "OK", failures);
  return failures ? 1 : 0;
}
