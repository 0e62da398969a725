// Particle.cpp -- reference: src/Particle.cpp.  Both classes call
// sh_particles_move (the interpolation, the cell clamp, the position update
// and the periodic wrap run on the device, bit-identical to the reference's
// double arithmetic); the host keeps only copies for position()/energy().
#include "Particle.H"

#include <atomic>
#include <cstdlib>
#include <iostream>

#include "spectral_hip.h"

#define ERROR(message)                                                                        \
    do {                                                                                      \
        std::cout << "ERROR in " << __func__ << " (" << __FILE__ << " " << __LINE__ << "): " \
                  << message << std::endl;                                                    \
        std::exit(1);                                                                         \
    } while (0)

#define CHECK(call)                                                                     \
    do {                                                                                \
        int _rc = (call);                                                               \
        if (_rc != 0) ERROR(#call << " failed (" << _rc << "): " << sh_last_error()); \
    } while (0)

namespace {

// process-wide tracer counter for Particle() (the reference's rand() stream);
// atomic because drivers construct and move tracers from OpenMP threads
// (exec/ns2.cpp:191-197)
std::atomic<long> g_next_tracer{0};

// device staging for single-tracer moves (position, velocity, energy), one
// per host thread so concurrent moves do not share it
struct Staging {
    double* p = nullptr;
    ~Staging() {
        if (p) sh_free(p);
    }
};
thread_local Staging t_one;

double* staging() {
    if (!t_one.p) CHECK(sh_malloc((void**)&t_one.p, 7 * sizeof(double)));
    return t_one.p;
}

}  // namespace

Particle::Particle() : m_v{0, 0, 0} {
    // tracer p of the sequence takes counters 3p .. 3p+2 of seed 1, the same
    // positions as Particles(np, 1)
    double* d = staging();
    CHECK(sh_particles_seed(d, 1, 1, g_next_tracer++, nullptr));
    CHECK(sh_memcpy_d2h(m_loc.data(), d, 3 * sizeof(double)));
}

void Particle::move(const Vector& velocity, double K) {
    std::array<INT, DIM> dims = velocity.getDims();
    double* d = staging();
    CHECK(sh_memcpy_h2d(d, m_loc.data(), 3 * sizeof(double)));
    CHECK(sh_particles_move(d, d + 3, d + 6, 1, velocity[0].deviceData(), velocity[1].deviceData(),
                            velocity[2].deviceData(), (int)dims[0], (int)dims[1], (int)dims[2], K, nullptr));
    double h[6];
    CHECK(sh_memcpy_d2h(h, d, 6 * sizeof(double)));
    for (int i = 0; i < DIM; i++) {
        m_loc[i] = h[i];
        m_v[i] = h[3 + i];
    }
}

double Particle::energy() {
    double e = 0;
    for (int d = 0; d < DIM; d++) e += m_v[d] * m_v[d];
    return .5 * e;
}

std::array<double, DIM> Particle::position() { return m_loc; }

Particles::Particles(INT np, uint64_t seed)
    : m_np(np), m_loc(nullptr), m_vel(nullptr), m_energy(nullptr), m_host_current(false) {
    if (np > 0) {
        CHECK(sh_malloc((void**)&m_loc, 3 * np * sizeof(double)));
        CHECK(sh_malloc((void**)&m_vel, 3 * np * sizeof(double)));
        CHECK(sh_malloc((void**)&m_energy, np * sizeof(double)));
        CHECK(sh_particles_seed(m_loc, (long)np, seed, 0, nullptr));
        CHECK(sh_memset(m_vel, 0, 3 * np * sizeof(double), nullptr));
        CHECK(sh_memset(m_energy, 0, np * sizeof(double), nullptr));
    }
    m_hloc.assign(3 * np, 0.0);
    m_henergy.assign(np, 0.0);
}

Particles::~Particles() {
    if (m_loc) sh_free(m_loc);
    if (m_vel) sh_free(m_vel);
    if (m_energy) sh_free(m_energy);
}

INT Particles::size() const { return m_np; }

void Particles::move(const Vector& velocity, double K) {
    if (m_np == 0) return;
    std::array<INT, DIM> dims = velocity.getDims();
    CHECK(sh_particles_move(m_loc, m_vel, m_energy, (long)m_np, velocity[0].deviceData(),
                            velocity[1].deviceData(), velocity[2].deviceData(), (int)dims[0], (int)dims[1],
                            (int)dims[2], K, nullptr));
    m_host_current = false;
}

void Particles::download() {
    if (m_host_current || m_np == 0) return;
    CHECK(sh_memcpy_d2h(m_hloc.data(), m_loc, 3 * m_np * sizeof(double)));
    CHECK(sh_memcpy_d2h(m_henergy.data(), m_energy, m_np * sizeof(double)));
    m_host_current = true;
}

std::array<double, DIM> Particles::position(INT i) {
    download();
    return {m_hloc[3 * i], m_hloc[3 * i + 1], m_hloc[3 * i + 2]};
}

double Particles::energy(INT i) {
    download();
    return m_henergy[i];
}

const double* Particles::deviceLocations() const { return m_loc; }
const double* Particles::deviceVelocities() const { return m_vel; }
const double* Particles::deviceEnergies() const { return m_energy; }
