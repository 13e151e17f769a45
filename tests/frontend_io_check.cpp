// Writes one msgpack frame and one .eigen.bin with the front-end encoders,
// and decodes the frame with the front-end reader (test helper, compiled by
// tests/test_frontends.py).
#include "front_common.hpp"

int main(int argc, char** argv) {
    if (argc != 3) return 2;
    // 3 boxes x 5 modes; box 1 empty; counts that exercise every integer width
    const int32_t H[15] = {0, 1, -1, 127, 128,  0, 0, 0, 0, 0,  -32, -33, 70000, -129, -40000};
    std::vector<double> temps = {300.0, 301.5, 299.25};
    std::string f = front::pack_frame(12.5, 3.25e-24, temps, 3, 5, H);
    // the reader used by dist_reader decodes what the writer wrote
    {
        front::Unpack u(f.data(), f.size());
        if (u.array() != 3 + 2 * 3 || u.number() != 12.5 || u.number() != 3.25e-24 || u.array() != 3) return 3;
        for (double t : temps)
            if (u.number() != t) return 3;
        for (int b = 0; b < 3; b++) {
            const int nel = (int)u.number();
            const uint32_t na = u.array();
            if (nel == 0) {
                if (na != 1 || u.number() != 42.0) return 3;
                continue;
            }
            for (uint32_t k = 0; k < na / 2; k++) {
                const int mode = (int)u.number();
                if (u.number() != (double)H[b * 5 + mode]) return 3;
            }
        }
    }
    std::ofstream o(argv[1], std::ios::binary);
    o << f.size() << "#" << f;
    f = front::pack_frame(17.5, 3.25e-24, temps, 3, 5, H);
    o << f.size() << "#" << f << "0#";
    o.close();
    std::vector<double> a(9);
    for (int i = 0; i < 9; i++) a[i] = 0.5 * i - 1.0;
    front::save_eigen_bin(argv[2], 3, a.data());
    int64_t rows = 0;
    auto b = front::load_eigen_bin(argv[2], rows);
    return rows == 3 && b == a ? 0 : 1;
}
