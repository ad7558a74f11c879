// split_operator.hpp -- header-only C++ facade over the C ABI (include/tdse.h)
// with the reference core's names and signatures, so reference-style call
// sites compile unchanged in shape:
//
//   ses::Grid1D / Grid3D                 core/src/grid.ixx:10-30
//   ses::Field1D / Field3D               core/src/field.ixx:11-66
//   ses::SplitOperator1D(g, V, dt)       core/src/propagator.ixx:21-64
//   ses::SplitOperator3D(g, V, dt, m)    core/src/propagator.ixx:66-134
//   ses::WavepacketSimulation            core/src/simulation.ixx:22-88
//   ses::ImaginaryTimePropagator1D/3D    core/src/imaginary_time.ixx:21-144
//   ses::DipoleDrive, ses::driven_step   core/src/drive.ixx:19-66
//   ses::MagneticPropagator3D            core/src/magnetic.ixx:19-80
//   ses::PeierlsLattice2D                core/src/lattice2d.ixx:27-275
//
// step(Field&, n) uploads the field, runs n Strang steps on the GPU and reads
// it back (the reference's in-place contract).  Device-resident stepping is
// tdse::Propagator.  Unsupported sizes throw std::invalid_argument like the
// reference (core/src/fft.ixx:42-44); device errors throw std::runtime_error.
#pragma once

#include <complex>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../tdse.h"

namespace tdse {

inline void check(int status) {
    if (status == TDSE_OK) return;
    std::string msg = tdse_last_error();
    if (msg.empty()) msg = tdse_status_string(status);
    if (status == TDSE_E_INVALID || status == TDSE_E_UNSUPPORTED) throw std::invalid_argument(msg);
    throw std::runtime_error(msg);
}

struct Grid1D {
    double xmin{};
    double xmax{};
    int n{};
    constexpr int size() const noexcept { return n; }
    constexpr double spacing() const noexcept { return (xmax - xmin) / n; }
    constexpr double coord(int i) const noexcept { return xmin + i * spacing(); }
};

struct Grid3D {
    Grid1D x{}, y{}, z{};
    constexpr int size() const noexcept { return x.n * y.n * z.n; }
    constexpr int flat(int i, int j, int k) const noexcept { return i + x.n * (j + y.n * k); }
    constexpr double cell_volume() const noexcept { return x.spacing() * y.spacing() * z.spacing(); }
};

struct Vec3d {
    double x{}, y{}, z{};
};

class Field1D {
public:
    explicit Field1D(Grid1D g) : grid_(g), data_(static_cast<std::size_t>(g.n)) {}
    int size() const noexcept { return grid_.n; }
    const Grid1D& grid() const noexcept { return grid_; }
    std::complex<double>& operator[](int i) noexcept { return data_[static_cast<std::size_t>(i)]; }
    const std::complex<double>& operator[](int i) const noexcept {
        return data_[static_cast<std::size_t>(i)];
    }
    std::vector<std::complex<double>>& data() noexcept { return data_; }
    const std::vector<std::complex<double>>& data() const noexcept { return data_; }

private:
    Grid1D grid_;
    std::vector<std::complex<double>> data_;
};

class Field3D {
public:
    explicit Field3D(Grid3D g) : grid_(g), data_(static_cast<std::size_t>(g.size())) {}
    int size() const noexcept { return grid_.size(); }
    const Grid3D& grid() const noexcept { return grid_; }
    std::complex<double>& operator()(int i, int j, int k) noexcept {
        return data_[static_cast<std::size_t>(grid_.flat(i, j, k))];
    }
    const std::complex<double>& operator()(int i, int j, int k) const noexcept {
        return data_[static_cast<std::size_t>(grid_.flat(i, j, k))];
    }
    std::vector<std::complex<double>>& data() noexcept { return data_; }
    const std::vector<std::complex<double>>& data() const noexcept { return data_; }

private:
    Grid3D grid_;
    std::vector<std::complex<double>> data_;
};

// A device and its HBM budget (0 = none).  One default context per process.
class Context {
public:
    explicit Context(int device = 0, std::uint64_t budget = 0) {
        tdse_ctx* c = nullptr;
        check(tdse_ctx_create(device, budget, &c));
        ctx_.reset(c);
    }
    tdse_ctx* get() const noexcept { return ctx_.get(); }
    static Context& instance() {
        static Context c;
        return c;
    }

private:
    struct Del {
        void operator()(tdse_ctx* c) const noexcept { tdse_ctx_destroy(c); }
    };
    std::unique_ptr<tdse_ctx, Del> ctx_;
};

enum class Precision : int { FP32 = TDSE_FP32, FP64 = TDSE_FP64 };

// Device-resident propagator (one plan): the state lives in HBM between calls.
class Propagator {
public:
    // bfield != 0: a uniform field along b_axis (set before the potential,
    // which then carries the diamagnetic term)
    Propagator(const Grid3D& g, const std::vector<double>& potential, double dt, double mass = 1.0,
               Precision prec = Precision::FP64, Context& ctx = Context::instance(),
               unsigned flags = 0u, double bfield = 0.0, int b_axis = 2) {
        const int n[3] = {g.x.n, g.y.n, g.z.n};
        const double lo[3] = {g.x.xmin, g.y.xmin, g.z.xmin};
        const double hi[3] = {g.x.xmax, g.y.xmax, g.z.xmax};
        if (static_cast<int>(potential.size()) != g.size())
            throw std::invalid_argument("potential size does not match the grid");
        tdse_plan* p = nullptr;
        check(tdse_plan_create(ctx.get(), n, lo, hi, dt, mass, static_cast<int>(prec), flags, &p));
        plan_.reset(p);
        size_ = g.size();
        if (bfield != 0.0) check(tdse_plan_set_magnetic(plan_.get(), bfield, b_axis));
        check(tdse_plan_set_potential(plan_.get(), potential.data()));
    }
    void set_mask(const std::vector<double>& m) { check(tdse_plan_set_mask(plan_.get(), m.data())); }
    void clear_mask() { check(tdse_plan_set_mask(plan_.get(), nullptr)); }
    void upload(const std::vector<std::complex<double>>& psi) {
        check(tdse_state_upload(plan_.get(), reinterpret_cast<const double*>(psi.data())));
    }
    void download(std::vector<std::complex<double>>& psi) const {
        psi.resize(static_cast<std::size_t>(size_));
        check(tdse_state_download(plan_.get(), reinterpret_cast<double*>(psi.data())));
    }
    void step(int nsteps) { check(tdse_step(plan_.get(), nsteps)); }
    void sync(int timeout_ms = -1) { check(tdse_plan_sync(plan_.get(), timeout_ms)); }
    tdse_observables observe(unsigned what = TDSE_OBS_NORM | TDSE_OBS_POSITION | TDSE_OBS_ENERGY) {
        tdse_observables o{};
        check(tdse_observe(plan_.get(), what, &o));
        return o;
    }
    tdse_plan* get() const noexcept { return plan_.get(); }

private:
    struct Del {
        void operator()(tdse_plan* p) const noexcept { tdse_plan_destroy(p); }
    };
    std::unique_ptr<tdse_plan, Del> plan_;
    int size_ = 0;
};

inline Grid3D as_grid3(const Grid1D& g) { return Grid3D{g, Grid1D{0.0, 1.0, 1}, Grid1D{0.0, 1.0, 1}}; }

// The reference propagator's FP64 phase tables (core/src/propagator.ixx:107-108),
// built on the host on first use from the potential the propagator was given
// (tdse_build_phase_tables, bit-identical to the reference's constructor).
// The GPU path does not read them: it keeps reduced kick angles and per-axis
// kinetic tables.
class PhaseTables {
public:
    PhaseTables(const Grid3D& g, const std::vector<double>& v, double dt, double mass)
        : g_(g), v_(v), dt_(dt), mass_(mass) {}
    const std::vector<std::complex<double>>& half_potential_phase() const {
        build();
        return half_;
    }
    const std::vector<std::complex<double>>& kinetic_phase() const {
        build();
        return kin_;
    }

private:
    void build() const {
        if (built_) return;
        const int n[3] = {g_.x.n, g_.y.n, g_.z.n};
        const double lo[3] = {g_.x.xmin, g_.y.xmin, g_.z.xmin};
        const double hi[3] = {g_.x.xmax, g_.y.xmax, g_.z.xmax};
        half_.resize(v_.size());
        kin_.resize(v_.size());
        check(tdse_build_phase_tables(n, lo, hi, v_.data(), dt_, mass_,
                                      reinterpret_cast<double*>(half_.data()),
                                      reinterpret_cast<double*>(kin_.data())));
        built_ = true;
    }
    Grid3D g_;
    std::vector<double> v_;
    double dt_, mass_;
    mutable bool built_ = false;
    mutable std::vector<std::complex<double>> half_, kin_;
};

// ses::SplitOperator1D (core/src/propagator.ixx:21-64)
class SplitOperator1D {
public:
    SplitOperator1D(const Grid1D& g, const std::vector<double>& potential, double dt,
                    Precision prec = Precision::FP64)
        : dt_(dt), prop_(std::make_unique<Propagator>(as_grid3(g), potential, dt, 1.0, prec)),
          tables_(as_grid3(g), potential, dt, 1.0) {}
    double dt() const noexcept { return dt_; }
    void step(Field1D& psi, int nsteps = 1) const {
        if (nsteps <= 0) return;
        prop_->upload(psi.data());
        prop_->step(nsteps);
        prop_->download(psi.data());
    }
    const std::vector<std::complex<double>>& half_potential_phase() const {
        return tables_.half_potential_phase();
    }
    const std::vector<std::complex<double>>& kinetic_phase() const { return tables_.kinetic_phase(); }
    Propagator& device() const noexcept { return *prop_; }

private:
    double dt_;
    std::unique_ptr<Propagator> prop_;
    PhaseTables tables_;
};

// ses::SplitOperator3D (core/src/propagator.ixx:66-134); 2D is nx x ny x 1.
class SplitOperator3D {
public:
    SplitOperator3D(const Grid3D& g, const std::vector<double>& potential, double dt,
                    double mass = 1.0, Precision prec = Precision::FP64)
        : dt_(dt), prop_(std::make_unique<Propagator>(g, potential, dt, mass, prec)),
          tables_(g, potential, dt, mass) {}
    double dt() const noexcept { return dt_; }
    void step(Field3D& psi, int nsteps = 1) const {
        if (nsteps <= 0) return;
        prop_->upload(psi.data());
        prop_->step(nsteps);
        prop_->download(psi.data());
    }
    // core/src/propagator.ixx:107-108 (FP64 host tables, built on first use)
    const std::vector<std::complex<double>>& half_potential_phase() const {
        return tables_.half_potential_phase();
    }
    const std::vector<std::complex<double>>& kinetic_phase() const { return tables_.kinetic_phase(); }
    // the device plan behind this propagator (driven_step sets its drive)
    Propagator& device() const noexcept { return *prop_; }

private:
    double dt_;
    std::unique_ptr<Propagator> prop_;
    PhaseTables tables_;
};

// ses::DipoleDrive (core/src/drive.ixx:19-23): V = amplitude (axis . r) cos(omega t)
struct DipoleDrive {
    Vec3d axis;
    double amplitude{};
    double omega{};
};

// ses::driven_step (core/src/drive.ixx:52-66): nsteps driven Strang steps of
// `prop` from time t0; the drive's half kicks ride in the kick sweeps.  The
// propagator's plan is left without a drive afterwards.
inline void driven_step(Field3D& psi, const SplitOperator3D& prop, const DipoleDrive& d, double t0,
                        int nsteps) {
    if (nsteps <= 0) return;
    Propagator& p = prop.device();
    const double axis[3] = {d.axis.x, d.axis.y, d.axis.z};
    check(tdse_plan_set_drive(p.get(), axis, d.amplitude, d.omega));
    check(tdse_plan_set_time(p.get(), t0));
    p.upload(psi.data());
    p.step(nsteps);
    p.download(psi.data());
    check(tdse_plan_set_drive(p.get(), axis, 0.0, 0.0));
}

// ses::MagneticPropagator3D (core/src/magnetic.ixx:19-80): uniform field
// `bfield` along `axis` (0 = x, 1 = y, 2 = z), symmetric gauge; each step is
// R(a) . Strang(V + B^2/8 |r_perp|^2) . R(a), a = (B/2)(dt/2).
class MagneticPropagator3D {
public:
    MagneticPropagator3D(const Grid3D& g, const std::vector<double>& v, double dt, double bfield,
                         int axis = 2, Precision prec = Precision::FP64)
        : dt_(dt), veff_(v),
          prop_(std::make_unique<Propagator>(g, v, dt, 1.0, prec, Context::instance(), 0u, bfield,
                                             axis)) {
        // host copy of the diamagnetic term the plan folded into its potential
        const double c = bfield * bfield / 8.0;
        if (c != 0.0)
            for (int k = 0; k < g.z.n; ++k)
                for (int j = 0; j < g.y.n; ++j)
                    for (int i = 0; i < g.x.n; ++i) {
                        const double r[3] = {g.x.coord(i), g.y.coord(j), g.z.coord(k)};
                        double perp = 0.0;
                        for (int a = 0; a < 3; ++a)
                            if (a != axis) perp += r[a] * r[a];
                        veff_[static_cast<std::size_t>(g.flat(i, j, k))] += c * perp;
                    }
    }
    double dt() const noexcept { return dt_; }
    const std::vector<double>& effective_potential() const noexcept { return veff_; }
    void step(Field3D& psi, int nsteps = 1) const {
        if (nsteps <= 0) return;
        prop_->upload(psi.data());
        prop_->step(nsteps);
        prop_->download(psi.data());
    }

private:
    double dt_;
    std::vector<double> veff_;
    std::unique_ptr<Propagator> prop_;
};

// ses::PeierlsLattice2D (core/src/lattice2d.ixx:27-275): finite-difference
// hopping with Peierls link phases on an n x m x 1 grid.  step propagates,
// relax relaxes in imaginary time (a second plan with the same links),
// energy is the lattice <H>.
class PeierlsLattice2D {
public:
    PeierlsLattice2D(const Grid3D& g, const std::vector<double>& potential, double dt,
                     Precision prec = Precision::FP64)
        : grid_(g), v_(potential), dt_(dt), prec_(prec),
          prop_(std::make_unique<Propagator>(g, potential, dt, 1.0, prec, Context::instance(),
                                             TDSE_FLAG_LATTICE)) {}
    double dt() const noexcept { return dt_; }
    void set_solenoid(double phi, double xs, double ys, bool cut_up = true) {
        links_ = {2, phi, xs, ys, cut_up};
        apply(*prop_);
        if (relaxer_) apply(*relaxer_);
    }
    void set_uniform_field(double b) {
        links_ = {1, b, 0.0, 0.0, true};
        apply(*prop_);
        if (relaxer_) apply(*relaxer_);
    }
    void step(Field3D& psi, int nsteps = 1) const { run(*prop_, psi, nsteps); }
    void relax(Field3D& psi, int nsteps = 1) const {
        if (!relaxer_) {
            relaxer_ = std::make_unique<Propagator>(grid_, v_, dt_, 1.0, prec_, Context::instance(),
                                                    TDSE_FLAG_LATTICE | TDSE_FLAG_IMAGINARY_TIME);
            apply(*relaxer_);
        }
        run(*relaxer_, psi, nsteps);
    }
    double energy(const Field3D& psi) const {
        prop_->upload(psi.data());
        return prop_->observe(TDSE_OBS_NORM | TDSE_OBS_ENERGY).energy;
    }

private:
    struct Links {
        int kind;  // 0 none, 1 uniform field, 2 solenoid
        double a, xs, ys;
        bool cut_up;
    };
    void apply(Propagator& p) const {
        if (links_.kind == 1) check(tdse_lattice_set_uniform_field(p.get(), links_.a));
        if (links_.kind == 2)
            check(tdse_lattice_set_solenoid(p.get(), links_.a, links_.xs, links_.ys,
                                            links_.cut_up ? 1 : 0));
    }
    static void run(Propagator& p, Field3D& psi, int nsteps) {
        if (nsteps <= 0) return;
        p.upload(psi.data());
        p.step(nsteps);
        p.download(psi.data());
    }
    Grid3D grid_;
    std::vector<double> v_;
    double dt_;
    Precision prec_;
    Links links_{0, 0.0, 0.0, 0.0, true};
    std::unique_ptr<Propagator> prop_;
    mutable std::unique_ptr<Propagator> relaxer_;
};

// ses::ImaginaryTimePropagator1D (core/src/imaginary_time.ixx:21-63): relax
// steps e^{-H dtau} with a renormalisation after each.
class ImaginaryTimePropagator1D {
public:
    ImaginaryTimePropagator1D(const Grid1D& g, const std::vector<double>& potential, double dtau,
                              Precision prec = Precision::FP64)
        : prop_(std::make_unique<Propagator>(as_grid3(g), potential, dtau, 1.0, prec,
                                             Context::instance(), TDSE_FLAG_IMAGINARY_TIME)) {}
    void relax(Field1D& psi, int nsteps) const {
        if (nsteps <= 0) return;
        prop_->upload(psi.data());
        prop_->step(nsteps);
        prop_->download(psi.data());
    }
    tdse_relax_stats stats() const {
        tdse_relax_stats s{};
        check(tdse_plan_relax_stats(prop_->get(), &s));
        return s;
    }

private:
    std::unique_ptr<Propagator> prop_;
};

// ses::ImaginaryTimePropagator3D (core/src/imaginary_time.ixx:65-144).
class ImaginaryTimePropagator3D {
public:
    ImaginaryTimePropagator3D(const Grid3D& g, const std::vector<double>& potential, double dtau,
                              double mass = 1.0, Precision prec = Precision::FP64)
        : grid_(g), dtau_(dtau), prec_(prec),
          prop_(std::make_unique<Propagator>(g, potential, dtau, mass, prec, Context::instance(),
                                             TDSE_FLAG_IMAGINARY_TIME)) {}
    void relax(Field3D& psi, int nsteps) const {
        if (nsteps <= 0) return;
        prop_->upload(psi.data());
        prop_->step(nsteps);
        prop_->download(psi.data());
    }
    // Gram-Schmidt deflation of `lower` each step -> the next state up.
    void relax_deflated(Field3D& psi, const std::vector<const Field3D*>& lower, int nsteps) const {
        std::vector<std::unique_ptr<Propagator>> held;
        std::vector<const tdse_plan*> plans;
        const std::vector<double> zero(static_cast<std::size_t>(grid_.size()), 0.0);
        for (const Field3D* q : lower) {
            held.push_back(std::make_unique<Propagator>(grid_, zero, dtau_, 1.0, prec_));
            held.back()->upload(q->data());
            plans.push_back(held.back()->get());
        }
        prop_->upload(psi.data());
        check(tdse_relax_deflated(prop_->get(), plans.data(), static_cast<int>(plans.size()),
                                  nsteps));
        prop_->download(psi.data());
    }
    tdse_relax_stats stats() const {
        tdse_relax_stats s{};
        check(tdse_plan_relax_stats(prop_->get(), &s));
        return s;
    }

private:
    Grid3D grid_;
    double dtau_;
    Precision prec_;
    std::unique_ptr<Propagator> prop_;
};

// core/src/wavepacket.ixx:27-46 (bit-identical builder in the library)
inline Field3D gaussian_wavepacket(const Grid3D& g, Vec3d r0, Vec3d sigma, Vec3d k0) {
    Field3D f{g};
    const int n[3] = {g.x.n, g.y.n, g.z.n};
    const double lo[3] = {g.x.xmin, g.y.xmin, g.z.xmin}, hi[3] = {g.x.xmax, g.y.xmax, g.z.xmax};
    const double r[3] = {r0.x, r0.y, r0.z}, s[3] = {sigma.x, sigma.y, sigma.z},
                 k[3] = {k0.x, k0.y, k0.z};
    check(tdse_build_gaussian(n, lo, hi, r, s, k, reinterpret_cast<double*>(f.data().data())));
    return f;
}

// ses::WavepacketSimulation (core/src/simulation.ixx:22-88): advance() and
// relax() keep the state on the GPU; psi() reads it back when it changed.
class WavepacketSimulation {
public:
    struct Config {
        Grid3D grid;
        std::vector<double> potential;
        Vec3d r0;
        Vec3d sigma;
        Vec3d k0;
        double dt;
    };
    explicit WavepacketSimulation(Config cfg, Precision prec = Precision::FP64)
        : grid_(cfg.grid), dt_(cfg.dt), prec_(prec), potential_(std::move(cfg.potential)),
          propagator_(grid_, potential_, cfg.dt, 1.0, prec), psi_(grid_) {
        psi_ = gaussian_wavepacket(grid_, cfg.r0, cfg.sigma, cfg.k0);
        propagator_.device().upload(psi_.data());
    }
    void advance(int nsteps) {
        propagator_.device().step(nsteps);
        steps_ += nsteps;
        stale_ = true;
    }
    // core/src/simulation.ixx:45-52: an imaginary-time relaxer keyed on dtau
    // (rebuilt when dtau changes); tau is not t, so time() does not move.  The
    // state goes to the relaxer's plan and back on the GPU.
    void relax(int nsteps, double dtau) {
        if (!relaxer_ || relaxer_dtau_ != dtau) {
            relaxer_ = std::make_unique<Propagator>(grid_, potential_, dtau, 1.0, prec_,
                                                    Context::instance(), TDSE_FLAG_IMAGINARY_TIME);
            relaxer_dtau_ = dtau;
        }
        if (nsteps <= 0) return;
        check(tdse_state_copy(relaxer_->get(), propagator_.device().get()));
        relaxer_->step(nsteps);
        check(tdse_state_copy(propagator_.device().get(), relaxer_->get()));
        stale_ = true;
    }
    double time() const noexcept { return steps_ * dt_; }
    double dt() const noexcept { return dt_; }
    const Grid3D& grid() const noexcept { return grid_; }
    const std::vector<double>& potential() const noexcept { return potential_; }
    // core/src/simulation.ixx:68: the state, read back from the GPU only when
    // a step changed it since the last read
    const Field3D& psi() const {
        if (stale_) {
            propagator_.device().download(psi_.data());
            stale_ = false;
        }
        return psi_;
    }
    // core/src/simulation.ixx:70
    const SplitOperator3D& propagator() const noexcept { return propagator_; }
    void set_psi(const Field3D& psi) {
        psi_ = psi;
        propagator_.device().upload(psi_.data());
        stale_ = false;
    }
    std::vector<double> density() const {
        const Field3D& f = psi();
        std::vector<double> d(f.data().size());
        for (std::size_t i = 0; i < d.size(); ++i) d[i] = std::norm(f.data()[i]);
        return d;
    }

private:
    Grid3D grid_;
    double dt_;
    Precision prec_;
    std::vector<double> potential_;
    SplitOperator3D propagator_;
    mutable Field3D psi_;
    long long steps_ = 0;
    mutable bool stale_ = false;
    std::unique_ptr<Propagator> relaxer_;
    double relaxer_dtau_ = 0.0;
};

}  // namespace tdse
