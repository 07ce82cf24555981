#!/bin/bash
# Writes philox_rocrand.json: Philox4x32-10 outputs computed by rocRAND's own
# philox4x32_10_engine::ten_rounds (ROCm's open-source rocrand headers), used to
# pin the oracle's Philox.  Host-only program; no GPU needed.
set -euo pipefail
HERE="$(cd "$(dirname "$0")" && pwd)"
TMP="$(mktemp -d)"
cat > "$TMP/gen.cpp" <<'CPP'
#include <rocrand/rocrand_philox4x32_10.h>
#include <cstdio>
#include <cstdint>
struct Eng : rocrand_device::philox4x32_10_engine {
  Eng() : rocrand_device::philox4x32_10_engine(0, 0, 0) {}
  uint4 run(uint4 c, uint2 k) { return ten_rounds(c, k); }
};
int main() {
  const uint32_t cases[][6] = {
    {0u, 0u, 0u, 0u, 0u, 0u}, {1u, 0u, 0u, 0u, 0u, 0u}, {0u, 1u, 2u, 3u, 4u, 5u},
    {0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu},
    {0x243f6a88u, 0x85a308d3u, 0x13198a2eu, 0x03707344u, 0xa4093822u, 0x299f31d0u},
    {12345u, 7u, 0u, 0u, 12345u, 0u}, {99u, 1000u, 3u, 0u, 0xdeadbeefu, 0x01234567u},
    {0x80000000u, 0x7fffffffu, 1u, 2u, 3u, 4u}, {42u, 42u, 42u, 42u, 42u, 42u}};
  Eng e;
  printf("[\n");
  const int n = sizeof(cases) / sizeof(cases[0]);
  for (int i = 0; i < n; ++i) {
    const uint32_t *c = cases[i];
    uint4 o = e.run(uint4{c[0], c[1], c[2], c[3]}, uint2{c[4], c[5]});
    printf("  {\"ctr\": [%u, %u, %u, %u], \"key\": [%u, %u], \"out\": [%u, %u, %u, %u]}%s\n",
           c[0], c[1], c[2], c[3], c[4], c[5], o.x, o.y, o.z, o.w, i + 1 < n ? "," : "");
  }
  printf("]\n");
  return 0;
}
CPP
/opt/rocm/bin/hipcc -O1 -o "$TMP/gen" "$TMP/gen.cpp"
"$TMP/gen" > "$HERE/philox_rocrand.json"
rm -rf "$TMP"
echo "wrote $HERE/philox_rocrand.json"
