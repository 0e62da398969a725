// Particle.H -- marker particles (src/Particle.H:7-20) on the device.
//
// Particle keeps the reference's per-tracer interface: constructed at a
// pseudo-random position in [0, 2 pi)^3, moved through a physical velocity
// field by trilinear interpolation, with energy() and position().  Each move
// is one device launch for one tracer, so a loop over many Particle objects
// (exec/ns2.cpp:191-204) is launch-bound; Particles holds a whole set in HBM
// and moves every tracer in one launch, with the same arithmetic.
//
// Positions come from splitmix64 (counter = 3 * tracer index + axis) rather
// than rand(): the reference's rand() sequence is platform-dependent.
#ifndef SPECTRAL_PARTICLE_H
#define SPECTRAL_PARTICLE_H

#include <array>
#include <cstdint>
#include <vector>

#include "DataObjects.H"

/// one tracer (src/Particle.H:7-20)
class Particle {
public:
    /// next position of the process-wide sequence (seed 1, one tracer per call)
    Particle();
    /// src/Particle.cpp:12-46 on the device.  Safe to call from several host
    /// threads at once when the velocity's device copy is current (as after
    /// Vector::inverse); a host-modified velocity must be uploaded first.
    void move(const Vector& velocity, double K);
    /// 0.5 |v|^2 of the last move (src/Particle.cpp:48-55)
    double energy();
    std::array<double, DIM> position();

private:
    std::array<double, DIM> m_loc;
    std::array<double, DIM> m_v;
};

/// a set of tracers resident in device memory, moved in one launch
class Particles {
public:
    Particles(INT np, uint64_t seed = 1);
    ~Particles();
    Particles(const Particles&) = delete;
    Particles& operator=(const Particles&) = delete;

    INT size() const;
    /// Particle::move for every tracer; velocity is the physical field
    void move(const Vector& velocity, double K);
    /// host copies, downloaded once after each move
    std::array<double, DIM> position(INT i);
    double energy(INT i);
    /// device arrays: positions and velocities [np][3], energies [np]
    const double* deviceLocations() const;
    const double* deviceVelocities() const;
    const double* deviceEnergies() const;

private:
    void download();

    INT m_np;
    double* m_loc;
    double* m_vel;
    double* m_energy;
    std::vector<double> m_hloc, m_henergy;
    bool m_host_current;
};

#endif
