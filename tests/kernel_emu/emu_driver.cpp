// Runs the library (compiled for the CPU emulation) on a case file and
// writes W, q, h, b after the requested number of steps.  TEST
// INFRASTRUCTURE ONLY.
//
// case file: the raw hugks_desc bytes (pointer fields ignored), then the
// arrays u[nv], v[nv] (2D), weight[nn], prim0[(dim+2)*ncell],
// accel_cell[2*ncell], accel_vface[2*(nx+1)*ny], accel_hface[2*nx*(ny+1)]
// (the last three in 2D only), all float64.
// output: int32 status, int32 steps, float64 simtime, dt, then
// w[(dim+2)*ncell], q[dim*ncell], h[ncell*nn], b[ncell*nn].
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "hugks.h"

static std::vector<double> rd(FILE* f, size_t n) {
    std::vector<double> v(n);
    if (n && fread(v.data(), sizeof(double), n, f) != n) {
        fprintf(stderr, "short case file\n");
        exit(2);
    }
    return v;
}

int main(int argc, char** argv) {
    if (argc < 4) {
        fprintf(stderr, "usage: emu_driver case.bin steps out.bin\n");
        return 2;
    }
    FILE* f = fopen(argv[1], "rb");
    if (!f) return 2;
    hugks_desc d;
    if (fread(&d, sizeof d, 1, f) != 1) return 2;
    const size_t ncell = (size_t)d.nx * d.ny;
    const size_t nn = d.dim == 2 ? (size_t)d.nv * d.nv : (size_t)d.nv;
    auto u = rd(f, d.nv);
    std::vector<double> v, ac, av, ah;
    if (d.dim == 2) v = rd(f, d.nv);
    auto w = rd(f, nn);
    auto p0 = rd(f, (d.dim + 2) * ncell);
    if (d.dim == 2) {
        ac = rd(f, 2 * ncell);
        av = rd(f, 2 * (size_t)(d.nx + 1) * d.ny);
        ah = rd(f, 2 * (size_t)d.nx * (d.ny + 1));
    }
    fclose(f);
    d.u = u.data();
    d.v = d.dim == 2 ? v.data() : nullptr;
    d.weight = w.data();
    d.prim0 = p0.data();
    d.accel_cell = d.dim == 2 ? ac.data() : nullptr;
    d.accel_vface = d.dim == 2 ? av.data() : nullptr;
    d.accel_hface = d.dim == 2 ? ah.data() : nullptr;
    hugks_handle* h = nullptr;
    int rc = hugks_create(&d, &h);
    if (rc != HUGKS_OK) {
        fprintf(stderr, "create: %d %s\n", rc, hugks_last_error(nullptr));
        return 3;
    }
    const int steps = atoi(argv[2]);
    hugks_step_info info{};
    rc = hugks_advance(h, steps, &info);
    if (rc != HUGKS_OK) fprintf(stderr, "advance: %d %s\n", rc, hugks_last_error(h));
    std::vector<double> W((d.dim + 2) * ncell), Q(d.dim * ncell), H(ncell * nn), B(ncell * nn);
    hugks_get_macro(h, W.data(), Q.data());
    hugks_get_f(h, H.data(), B.data());
    FILE* o = fopen(argv[3], "wb");
    int32_t hdr[2] = {rc, info.steps};
    fwrite(hdr, sizeof hdr, 1, o);
    fwrite(&info.simtime, sizeof(double), 1, o);
    fwrite(&info.dt, sizeof(double), 1, o);
    fwrite(W.data(), sizeof(double), W.size(), o);
    fwrite(Q.data(), sizeof(double), Q.size(), o);
    fwrite(H.data(), sizeof(double), H.size(), o);
    fwrite(B.data(), sizeof(double), B.size(), o);
    fclose(o);
    hugks_destroy(h);
    return rc == HUGKS_OK ? 0 : 1;
}
