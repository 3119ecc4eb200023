// Host-side check of the pass layout's index maps (rc_device.hpp): for each
// image size, order_to_pixel is a bijection of [0, w*h) and pixel_to_order
// its inverse; for each (pixels, samples per pass), slot -> (tile-order
// index, sample) is a bijection of [0, pixels * samples).  Exit status 0 when
// every check holds.  Built and run on the CPU by tests/test_pixel_order.py.
#include <cstdio>
#include <vector>

#include "rc_device.hpp"

using namespace rcd;

int main() {
    const unsigned sizes[][2] = {{1, 1},     {7, 5},     {8, 8},     {9, 17},   {64, 64},   {65, 130},  {1920, 1080},
                                 {256, 256}, {128, 128}, {1024, 1024}, {512, 512}, {1000, 3}, {3, 1000}, {800, 600},
                                 {96, 54},   {32, 32},   {48, 48},   {16, 16},  {200, 65}};
    int bad = 0;
    for (const auto& sz : sizes) {
        const unsigned w = sz[0], h = sz[1], np = w * h;
        const PixelOrder o = make_order(w, h);
        std::vector<unsigned char> seen(np, 0);
        for (unsigned q = 0; q < np; ++q) {
            const unsigned p = order_to_pixel(o, q);
            if (p >= np || seen[p] || pixel_to_order(o, p) != q) {
                std::printf("pixel order %ux%u fails at q=%u (p=%u)\n", w, h, q, p);
                ++bad;
                break;
            }
            seen[p] = 1;
        }
        for (unsigned spp : {1u, 3u, 64u}) {
            const unsigned n = np * spp;
            if (n > (1u << 24)) continue;
            std::vector<unsigned char> hit(n, 0);
            for (unsigned slot = 0; slot < n; ++slot) {
                const unsigned q = slot_index(slot, np, spp), j = slot_sample_in_pass(slot, np, spp);
                if (q >= np || j >= spp || hit[size_t(j) * np + q]) {
                    std::printf("slot map %u pixels x %u fails at slot %u\n", np, spp, slot);
                    ++bad;
                    break;
                }
                hit[size_t(j) * np + q] = 1;
            }
        }
    }
    std::printf("%d failures\n", bad);
    return bad != 0;
}
