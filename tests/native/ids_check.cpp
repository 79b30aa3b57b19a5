// Host check of the engine's voxel-id helpers (csrc/lsseg_device.hpp) against
// brute force over every voxel of a few tile geometries: the id is a bijection
// onto [0, N), decode inverts it, id_of_tau / tau_of_id convert to and from
// the reference's tiled index (Coordinates.cu:65-96), and the face neighbours
// and in-volume bits equal the clamp-to-self neighbours of the coordinates
// (Coordinates.cu:157-184).  Prints "ok" and exits 0 when every check holds.
#include "lsseg_device.hpp"

#include <cstdio>
#include <vector>

using namespace lsseg;

int main()
{
    long bad = 0;
    const uint32_t dims[][3] = {{3, 2, 2}, {1, 1, 1}, {2, 3, 4}, {5, 1, 2}};
    for (const auto& d : dims) {
        Geom g{};
        g.TX = d[0];
        g.TY = d[1];
        g.TZ = d[2];
        g.X = g.TX * 16;
        g.Y = g.TY * 16;
        g.Z = g.TZ * 16;
        g.sTY = g.TX * 4096u;
        g.sTZ = g.TX * g.TY * 4096u;
        g.N = (uint64_t)g.X * g.Y * g.Z;
        g.dTX = make_fastdiv(g.TX);
        g.dTXY = make_fastdiv(g.TX * g.TY);
        std::vector<char> seen(g.N, 0);
        const int off[7][3] = {{0, 0, -1}, {0, -1, 0}, {-1, 0, 0}, {0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
        for (uint32_t z = 0; z < g.Z; ++z)
            for (uint32_t y = 0; y < g.Y; ++y)
                for (uint32_t x = 0; x < g.X; ++x) {
                    const uint32_t id = id_of(g, x, y, z);
                    if (id >= g.N || seen[id]) {
                        ++bad;
                        continue;
                    }
                    seen[id] = 1;
                    uint32_t a, b, c;
                    decode_xyz(g, id, &a, &b, &c);
                    bad += (a != x || b != y || c != z);
                    const uint32_t tau = (((z >> 4) * g.TY + (y >> 4)) * g.TX + (x >> 4)) * 4096u +
                                         ((z & 15u) * 16u + (y & 15u)) * 16u + (x & 15u);
                    bad += id_of_tau(tau) != id;
                    bad += tau_of_id(id) != tau;
                    uint32_t nb[7], inb;
                    neighbors(g, id, nb, &inb);
                    for (int e = 0; e < 7; ++e) {
                        const long nx = (long)x + off[e][0], ny = (long)y + off[e][1], nz = (long)z + off[e][2];
                        const bool in = nx >= 0 && ny >= 0 && nz >= 0 && nx < (long)g.X && ny < (long)g.Y &&
                                        nz < (long)g.Z;
                        const uint32_t want = in ? id_of(g, (uint32_t)nx, (uint32_t)ny, (uint32_t)nz) : id;
                        bad += nb[e] != want;
                        if (e != 3) bad += (((inb >> e) & 1u) != 0) != in;
                        if (in) bad += neighbor_in(g, id, (uint32_t)e) != want;
                    }
                }
        printf("tiles %ux%ux%u: %llu voxels, failures %ld\n", g.TX, g.TY, g.TZ, (unsigned long long)g.N, bad);
    }
    if (bad) return 1;
    printf("ok\n");
    return 0;
}
