// ref_random_driver.cpp - prints the first draws of the reference's Random namespace
// (src/random.cpp, compiled from the reference tree by oracle/Makefile `ref`).
// Usage: ref_random <seed_a> <seed_b> <count>
// Output: <count> lines "u <hex of uniform bits>" then <count> lines of normals, %.17g.
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "random.h"

int main(int argc, char** argv) {
    if (argc != 4) {
        std::fprintf(stderr, "usage: %s seed_a seed_b count\n", argv[0]);
        return 2;
    }
    const unsigned a = (unsigned)std::strtoul(argv[1], nullptr, 10);
    const unsigned b = (unsigned)std::strtoul(argv[2], nullptr, 10);
    const int count = std::atoi(argv[3]);
    Random::seed(a, b);
    for (int i = 0; i < count; ++i) std::printf("u %.17g\n", Random::rand());
    Random::seed(a, b);
    for (int i = 0; i < count; ++i) std::printf("n %.17g\n", Random::normal());
    return 0;
}
