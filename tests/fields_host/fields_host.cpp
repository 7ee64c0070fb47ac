// Host driver for the library's field arithmetic (stwo-zig_amd/csrc/fields.hpp): reads
// "q a0 b0 a1 b1 c0 d0 c1 d1", "c a b c d" or "i a" lines and prints the QM31 product, the
// CM31 product or the M31 inverse, so that a test can compare them with an independent
// big-integer restatement of the field laws.
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstring>
#include "../../stwo-zig_amd/csrc/fields.hpp"
using namespace sr;
int main() {
    char op[4];
    while (scanf("%3s", op) == 1) {
        if (!strcmp(op, "q")) {
            uint32_t v[8];
            for (int k = 0; k < 8; k++)
                if (scanf("%u", &v[k]) != 1) return 1;
            const qm31 r = qmul({{v[0], v[1]}, {v[2], v[3]}}, {{v[4], v[5]}, {v[6], v[7]}});
            printf("%u %u %u %u\n", r.c0.a, r.c0.b, r.c1.a, r.c1.b);
        } else if (!strcmp(op, "c")) {
            uint32_t v[4];
            for (int k = 0; k < 4; k++)
                if (scanf("%u", &v[k]) != 1) return 1;
            const cm31 r = cmul({v[0], v[1]}, {v[2], v[3]});
            printf("%u %u\n", r.a, r.b);
        } else if (!strcmp(op, "i")) {
            uint32_t a;
            if (scanf("%u", &a) != 1) return 1;
            printf("%u\n", minv(a));
        } else {
            return 1;
        }
    }
    return 0;
}
