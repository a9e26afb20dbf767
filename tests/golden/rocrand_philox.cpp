// Prints rocRAND's Philox4x32-10 outputs (the MIT-licensed header in /opt/rocm) for a few
// seeds; tests/test_oracle.py compares them with the oracle's restatement.
// rocrand_init(seed, 0, 0) sets key = {seed_lo, seed_hi} and counter = 0; each rocrand4()
// call returns Philox(counter, key) and increments counter word 0.
#include <rocrand/rocrand_philox4x32_10.h>
#include <cstdio>
int main() {
    const unsigned long long seeds[] = {0ull, 1ull, 0x5eedull, 0x0123456789abcdefull, ~0ull};
    for (unsigned long long s : seeds) {
        rocrand_state_philox4x32_10 st;
        rocrand_init(s, 0, 0, &st);
        for (int i = 0; i < 3; i++) {
            uint4 v = rocrand4(&st);
            std::printf("%llu %d %u %u %u %u\n", s, i, v.x, v.y, v.z, v.w);
        }
    }
    return 0;
}
