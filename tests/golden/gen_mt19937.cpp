// Writes the seeded random particle sets of the reference's randomized grid
// and neighbour tests as raw little-endian float32 (x, y, z) triples:
//
//   grid_8193.f32     mt19937(0x5A17); x in [-12, 13), y in [-9, 11),
//                     z in [-7, 10) (test/test_spatial_grid.cu:649-672)
//   neighbors_257.f32 mt19937(12345); x, y, z in [0, 5)
//                     (test/test_neighbors.cu:342-353)
//
// The reference draws the three coordinates inside one call,
// position(d(e), d(e), d(e)), whose argument evaluation order C++ leaves to
// the compiler.  The same call expression is used here, so g++ (the host
// compiler of the reference's CUDA build) draws in the same order.
// Build and run through make_golden.py.
#include <cstdio>
#include <random>
#include <string>
#include <vector>

struct P {
    float x, y, z;
};

static P position(float x, float y, float z) { return P{x, y, z}; }

static void write(const std::string& path, const std::vector<P>& v) {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return;
    std::fwrite(v.data(), sizeof(P), v.size(), f);
    std::fclose(f);
}

int main(int argc, char** argv) {
    const std::string dir = argc > 1 ? argv[1] : ".";
    {
        std::mt19937 randomEngine(0x5A17U);
        std::uniform_real_distribution<float> xDistribution(-12.0f, 13.0f);
        std::uniform_real_distribution<float> yDistribution(-9.0f, 11.0f);
        std::uniform_real_distribution<float> zDistribution(-7.0f, 10.0f);
        std::vector<P> out;
        for (int i = 0; i < 8193; ++i)
            out.push_back(position(xDistribution(randomEngine), yDistribution(randomEngine),
                                   zDistribution(randomEngine)));
        write(dir + "/grid_8193.f32", out);
    }
    {
        std::mt19937 randomEngine(12345);
        std::uniform_real_distribution<float> distribution(0.0f, 5.0f);
        std::vector<P> out;
        for (int i = 0; i < 257; ++i)
            out.push_back(position(distribution(randomEngine), distribution(randomEngine),
                                   distribution(randomEngine)));
        write(dir + "/neighbors_257.f32", out);
    }
    return 0;
}
