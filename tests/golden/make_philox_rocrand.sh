#!/bin/sh
# Writes philox_rocrand.json: Philox4x32-10 blocks from rocRAND's own header implementation (host path of
# rocrand_philox4x32_10.h), an independent library routine to check the oracle's Philox against.
# rocrand_init(seed, subsequence, 0): key = {seed lo, seed hi}, counter = {0, 0, subsequence lo, hi};
# each rocrand() call returns the next word of the block, the counter incremented per block.
set -e
cd "$(dirname "$0")"
tmp=$(mktemp -d)
cat > "$tmp/p.cpp" <<'CPP'
#include <rocrand/rocrand_kernel.h>
#include <cstdio>
int main() {
  const unsigned long long seeds[3] = {0ull, 0x299f31d0a4093822ull, 0xdeadbeefcafef00dull};
  const unsigned long long subs[3] = {0ull, 1ull, 0x100000003ull};
  printf("[\n");
  int first = 1;
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) {
      rocrand_state_philox4x32_10 st;
      rocrand_init(seeds[a], subs[b], 0, &st);
      for (unsigned blk = 0; blk < 2; ++blk) {
        unsigned w[4];
        for (int i = 0; i < 4; ++i) w[i] = rocrand(&st);
        printf("%s  {\"seed\": %llu, \"ctr\": [%u, 0, %llu, %llu], \"out\": [%u, %u, %u, %u]}", first ? "" : ",\n",
               seeds[a], blk, subs[b] & 0xffffffffull, subs[b] >> 32, w[0], w[1], w[2], w[3]);
        first = 0;
      }
    }
  printf("\n]\n");
  return 0;
}
CPP
/opt/rocm/bin/hipcc -O1 "$tmp/p.cpp" -o "$tmp/p"
"$tmp/p" > philox_rocrand.json
rm -rf "$tmp"
