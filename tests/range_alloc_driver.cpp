// Drives the product allocator (pyrit_amd/csrc/range_alloc.h) from a script
// on stdin, one operation per line, for tests/test_range_alloc.py:
//   init C | alloc S | release O S  ->  prints the offset of each alloc (-1
//   when full) and, after every operation, "used <n> ranges <k>"
#include <stdio.h>
#include <string.h>

#include "range_alloc.h"

int main() {
    pyr::RangeAlloc ra(0);
    char op[16];
    unsigned long long a = 0, b = 0;
    while (scanf("%15s", op) == 1) {
        if (!strcmp(op, "init")) {
            if (scanf("%llu", &a) != 1) return 2;
            ra = pyr::RangeAlloc(a);
        } else if (!strcmp(op, "alloc")) {
            if (scanf("%llu", &a) != 1) return 2;
            printf("off %lld\n", (long long)ra.alloc(a));
        } else if (!strcmp(op, "release")) {
            if (scanf("%llu %llu", &a, &b) != 2) return 2;
            ra.release(a, b);
        } else {
            return 2;
        }
        printf("used %llu ranges %zu\n", (unsigned long long)ra.used(), ra.free_ranges());
    }
    return 0;
}
