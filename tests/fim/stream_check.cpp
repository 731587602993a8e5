// Host replay of the FIM spectrum kernel's walk of the random stream
// (fim_stream::walk_item, the function the kernel calls) against the
// sequential yarn2 stream (Yarn2::operator()) and against Yarn2::discard.
// Exit status 0 when every value matches and every point (line, i < h) of
// the half spectrum is visited exactly once.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "fim_stream.hpp"

using namespace scarf;

static int check_grid(const fim_stream::Tables &t, int seed, int n1, int n2, int n3, int per_item)
{
    const int h = n1 / 2 + 1;
    const long lines = long(n2) * n3;
    const long total = long(n1) * lines;
    std::vector<uint32_t> seq(total);
    Yarn2 y;
    y.seed(static_cast<unsigned long>(seed));
    for (long n = 0; n < total; ++n) seq[n] = static_cast<uint32_t>(y());
    uint32_t s0[2], line_step[4];
    fim_stream::position0(seed, s0);
    fim_stream::power(uint64_t(n1), line_step);
    std::vector<int> seen(size_t(h) * lines, 0);
    long bad = 0;
    const long items = (lines + per_item - 1) / per_item;
    for (long item = 0; item < items; ++item) {
        const long l0 = item * per_item, l1 = std::min<long>(l0 + per_item, lines);
        for (int lane = 0; lane < 64; ++lane)
            fim_stream::walk_item(t.pw, &t, s0, line_step, n1, h, l0, l1, lane, [&](long line, int i, uint32_t v) {
                seen[size_t(i) + size_t(h) * line] += 1;
                if (v != seq[size_t(i) + size_t(n1) * line]) ++bad;
            });
    }
    long miss = 0;
    for (int c : seen)
        if (c != 1) ++miss;
    std::printf("grid %d x %d x %d seed %d lines/item %d: %ld mismatches, %ld points not visited once\n", n1, n2, n3,
                seed, per_item, bad, miss);
    return bad == 0 && miss == 0 ? 0 : 1;
}

// far positions: the jump table against Yarn2::discard (matrix powers by %)
static int check_far(const fim_stream::Tables &t, int seed)
{
    int fails = 0;
    const uint64_t pos[] = {0, 1, 63, 64, 65, 1000003, (1ull << 31) - 1, (1ull << 31) + 12345, 8589934597ull,
                            (1ull << 40) - 3};
    uint32_t s0[2];
    fim_stream::position0(seed, s0);
    for (uint64_t n : pos) {
        uint32_t r0 = s0[0], r1 = s0[1];
        fim_stream::jump(t.jump, n, r0, r1);
        Yarn2 y;
        y.seed(static_cast<unsigned long>(seed));
        y.discard(n);
        const uint32_t want = static_cast<uint32_t>(y());
        const uint32_t got = fim_stream::value(t.pw, r0);
        if (want != got) {
            std::printf("position %llu seed %d: %u != %u\n", (unsigned long long)n, seed, got, want);
            ++fails;
        }
    }
    std::printf("far positions seed %d: %d mismatches\n", seed, fails);
    return fails;
}

int main()
{
    static fim_stream::Tables t;
    fim_stream::make_tables(t);
    int rc = 0;
    rc |= check_grid(t, 1235, 8, 6, 4, 8);
    rc |= check_grid(t, 7, 130, 10, 6, 8);    // h = 66: two points per lane for lanes 0, 1
    rc |= check_grid(t, -3, 300, 14, 1, 8);   // 2D, negative seed
    rc |= check_grid(t, 42, 2, 4, 2, 3);      // smallest even grid, ragged last item
    rc |= check_grid(t, 99, 258, 20, 18, 7);
    rc |= check_far(t, 1234) ? 1 : 0;
    rc |= check_far(t, 0) ? 1 : 0;
    return rc;
}
