// Reference-style call site compiled against the C++ facade: the shape of
// core/src/simulation.ixx (advance) and SplitOperator3D::step.  Writes the
// stepped state as raw complex128 to argv[1]; exits 77 without a GPU (the
// reference harness's skip code, solver/src/vkcheck_main.cpp:3824-3826).
#include <cmath>
#include <cstdio>
#include <fstream>
#include <stdexcept>
#include <vector>

#include "tdse/split_operator.hpp"

int main(int argc, char** argv) {
    const tdse::Grid3D g{tdse::Grid1D{-8.0, 8.0, 32}, tdse::Grid1D{-8.0, 8.0, 32},
                         tdse::Grid1D{-4.0, 4.0, 16}};
    std::vector<double> v(static_cast<std::size_t>(g.size()));
    for (int k = 0; k < g.z.n; ++k)
        for (int j = 0; j < g.y.n; ++j)
            for (int i = 0; i < g.x.n; ++i) {
                const double x = g.x.coord(i), y = g.y.coord(j), z = g.z.coord(k);
                v[static_cast<std::size_t>(g.flat(i, j, k))] = 0.5 * 0.25 * (x * x + y * y + z * z);
            }
    try {
        tdse::WavepacketSimulation sim({g, v, {1.0, 0.0, 0.0}, {1.0, 1.0, 1.0}, {0.0, 0.5, 0.0}, 0.02});
        sim.advance(7);
        sim.advance(5);
        tdse::Field3D psi = sim.psi();
        // step(Field&) on a second propagator from the same start: must agree
        tdse::SplitOperator3D op{g, v, 0.02};
        tdse::Field3D f = tdse::gaussian_wavepacket(g, {1.0, 0.0, 0.0}, {1.0, 1.0, 1.0}, {0.0, 0.5, 0.0});
        op.step(f, 12);
        double d = 0.0;
        for (std::size_t i = 0; i < f.data().size(); ++i) d = std::max(d, std::abs(f.data()[i] - psi.data()[i]));
        std::printf("facade: t=%.3f max|advance - step| = %.3e\n", sim.time(), d);
        try {
            tdse::SplitOperator3D bad{tdse::Grid3D{tdse::Grid1D{0, 1, 7}, tdse::Grid1D{0, 1, 8}, tdse::Grid1D{0, 1, 8}},
                                      std::vector<double>(7 * 64, 0.0), 0.1};
            std::printf("facade: size 7 accepted (wrong)\n");
            return 1;
        } catch (const std::invalid_argument&) {
        }
        // ImaginaryTimePropagator1D::relax (imaginary_time_test.cpp shape):
        // the harmonic ground state, E0 = 0.5 from the norm decay of a step
        const tdse::Grid1D g1{-16.0, 16.0, 512};
        std::vector<double> v1(512);
        for (int i = 0; i < 512; ++i) v1[static_cast<std::size_t>(i)] = 0.5 * g1.coord(i) * g1.coord(i);
        tdse::Field1D p1{g1};
        for (int i = 0; i < 512; ++i) {
            const double x = g1.coord(i) - 1.0;
            p1[i] = std::complex<double>{std::exp(-x * x / 16.0), 0.0};
        }
        tdse::ImaginaryTimePropagator1D relax{g1, v1, 0.005};
        relax.relax(p1, 4000);
        const double e0 = relax.stats().energy;
        std::printf("facade: relaxed E0 = %.8f\n", e0);
        if (std::abs(e0 - 0.5) > 1e-4) return 1;
        if (argc > 1) {
            std::ofstream out(argv[1], std::ios::binary);
            out.write(reinterpret_cast<const char*>(psi.data().data()),
                      static_cast<std::streamsize>(psi.data().size() * sizeof(std::complex<double>)));
        }
        return d < 1e-13 ? 0 : 1;
    } catch (const std::runtime_error& e) {
        std::printf("facade: %s\n", e.what());
        return 77;
    }
}
