// grasph2/initial_conditions.hpp -- GraSPH2's initial-condition generators.
// Host code only; no device calls.
//
//   InitGenerator<hostParticleType>   src/initialConditions/InitGenerator.h:36-155
//   ps::ParticleSource<P, Derived>    src/initialConditions/ParticleSource.h:49-203
//     (move, rotate, addPositionalNoise, addAngularVelocity,
//      addLinearVelocity, addRandomVelocity, setConstant<Attrib>)
//   ps::UniformSphere                 src/initialConditions/particleSources/UniformSphere.{h,cu}
//   ps::PlummerSphere                 src/initialConditions/particleSources/PlummerSphere.{h,cu}
//   (ps::TextFile, the snapshot reader, lives in grasph2/storage.hpp)
//
// Semantics kept from the reference:
//  - a source yields particle id by generating it and then applying its
//    modifiers in the order they were added; everything else of the
//    particle starts at zero;
//  - every modifier also transforms the source's "desired" centre of mass and
//    its velocity, starting from the origin at rest; generate() then, per
//    source and on request, shifts positions / velocities by the difference
//    between desired and actual mass-weighted mean (sums in f1_t, in
//    particle order);
//  - rotate() builds R = cos(a) I + (1 - cos(a)) u u^T + sin(a) [u]x for the
//    normalised axis u, applied as R * pos and R * vel (a right-handed
//    rotation by a about u);
//  - random draws: std::default_random_engine seeded by the caller (a
//    std::random_device value by default), std::uniform_real_distribution
//    of f1_t; three-component noise is drawn x, y, z;
//  - UniformSphere: mass M/N, density rho0, a point uniform in the ball;
//  - PlummerSphere: mass M/N, density rho0, a direction uniform on the unit
//    sphere and a radius a / sqrt(u^(-2/3) - 1) (the inverse of the Plummer
//    mass profile m(r)/M = (1 + a^2/r^2)^(-3/2)), redrawn while above the
//    cut-off; addRandomPlummerVelocity draws x by rejection from
//    g(x) = x^2 (1 - x^2)^3.5 (x ~ U(0,1), y ~ U(0,0.1), accept y <= g(x))
//    and adds x * sqrt(2 G M / sqrt(r^2 + a^2)) in a uniform direction, with
//    r the particle's distance from the origin when the modifier runs.
//
// The reference takes its points on the unit sphere and in the ball from
// mpUtils (mpu::randSphereShell, mpu::randUniformSphere; mpUtils v0.12.0,
// not part of its source tree).  Here they come from the inverse-CDF
// construction: azimuth 2 pi u1, cosine of the polar angle 1 - 2 u2
// (Archimedes' hat-box theorem), ball radius R u3^(1/3).  The distributions
// are the reference's; single draws do not reproduce a reference run.
// Two deliberate differences: every random modifier owns its engine (the
// reference keeps one function-static engine per modifier kind and source
// type, so the seed of a second such modifier is ignored), and the two or
// three draws of a point are taken in a fixed order (the reference passes
// them as arguments of one call, whose evaluation order C++ leaves open).
// A PlummerSphere with a cut-off radius <= 0 is rejected (the reference's
// redraw loop would never end).
#ifndef GRASPH2_INITIAL_CONDITIONS_HPP
#define GRASPH2_INITIAL_CONDITIONS_HPP

#include <algorithm>
#include <cmath>
#include <functional>
#include <random>
#include <stdexcept>
#include <type_traits>
#include <utility>
#include <vector>

#include "particles.hpp"

namespace ps {

// every attribute of the reference (src/particles/Particles.h full_particle)
using full_particle = Particle<POS, MASS, VEL, ACC, BALSARA, XVEL, DENSITY, DENSITY_DT, DSTRESS, DSTRESS_DT, MAXVSIG>;

namespace detail {
// tells particle sources apart from other types (ParticleSource.h:32-37)
class ParticleSourceBaseFlag {
};

inline unsigned int random_seed() { return std::random_device{}(); }

// direction uniform on the unit sphere from two U[0,1) draws
inline f3_t unit_direction(f1_t u1, f1_t u2)
{
    const f1_t phi = f1_t(6.283185307179586) * u1;
    const f1_t cz = f1_t(1) - f1_t(2) * u2;
    const f1_t sz = std::sqrt(std::max(f1_t(0), f1_t(1) - cz * cz));
    return {sz * std::cos(phi), sz * std::sin(phi), cz};
}

// point uniform in the ball of the given radius from three U[0,1) draws
inline f3_t ball_point(f1_t u1, f1_t u2, f1_t u3, f1_t radius)
{
    return unit_direction(u1, u2) * (radius * std::cbrt(u3));
}
}  // namespace detail

template <typename particleType, typename Derived>
class ParticleSource : public detail::ParticleSourceBaseFlag {
public:
    virtual ~ParticleSource() = default;

    size_t getNumParticles() const { return m_numberOfParticles; }
    f3_t getDesiredCOM() const { return m_desiredCOM; }
    f3_t getDesiredVelOfCOM() const { return m_desiredVelOfCOM; }

    // shift every particle by position
    template <typename vecType>
    Derived& move(vecType position)
    {
        const f3_t d{position.x, position.y, position.z};
        m_operations.push_back([d](full_particle& p) { p.pos += d; });
        m_desiredCOM += d;
        return self();
    }

    // rotate positions and velocities by angle (radians) about axis
    Derived& rotate(f3_t axis, f1_t angle)
    {
        const f3_t u = normalize(axis);
        const f1_t c = std::cos(angle), s = std::sin(angle);
        const m3_t uu(u.x * u.x, u.x * u.y, u.x * u.z, u.x * u.y, u.y * u.y, u.y * u.z, u.x * u.z, u.z * u.y,
                      u.z * u.z);
        const m3_t ux(0, -u.z, u.y, u.z, 0, -u.x, -u.y, u.x, 0);
        const m3_t R = c * m3_t(1.0f) + (1 - c) * uu + s * ux;
        m_operations.push_back([R](full_particle& p) {
            p.pos = R * p.pos;
            p.vel = R * p.vel;
        });
        m_desiredCOM = R * m_desiredCOM;
        m_desiredVelOfCOM = R * m_desiredVelOfCOM;
        return self();
    }

    // add U(-strength, strength) to every position component
    Derived& addPositionalNoise(f1_t strength, unsigned int seed = detail::random_seed())
    {
        m_operations.push_back([rng = std::default_random_engine(seed),
                                dist = std::uniform_real_distribution<f1_t>(-strength, strength)](
                                   full_particle& p) mutable { p.pos += f3_t{dist(rng), dist(rng), dist(rng)}; });
        return self();
    }

    // add omega x pos to every velocity
    Derived& addAngularVelocity(f3_t omega)
    {
        m_operations.push_back([omega](full_particle& p) { p.vel += cross(omega, p.pos); });
        m_desiredVelOfCOM += cross(omega, m_desiredCOM);
        return self();
    }

    Derived& addLinearVelocity(f3_t v)
    {
        m_operations.push_back([v](full_particle& p) { p.vel += v; });
        m_desiredVelOfCOM += v;
        return self();
    }

    // add U(-strength, strength) to every velocity component
    Derived& addRandomVelocity(f1_t strength, unsigned int seed = detail::random_seed())
    {
        m_operations.push_back([rng = std::default_random_engine(seed),
                                dist = std::uniform_real_distribution<f1_t>(-strength, strength)](
                                   full_particle& p) mutable { p.vel += f3_t{dist(rng), dist(rng), dist(rng)}; });
        return self();
    }

    // set attribute Attrib of every particle to v
    template <typename Attrib>
    Derived& setConstant(typename Attrib::type v)
    {
        m_operations.push_back([v](full_particle& p) { p.template setAttribute<Attrib>(v); });
        return self();
    }

    // generate particle id (< getNumParticles()) and store it at index pos
    template <typename particleBufferType>
    void operator()(particleBufferType& particles, size_t id, size_t pos)
    {
        full_particle p{};
        p = generateParticle(id);
        for (const auto& op : m_operations) op(p);
        particles.storeParticle(pos, p);
    }

protected:
    size_t m_numberOfParticles{0};
    f3_t m_desiredCOM{0, 0, 0};
    f3_t m_desiredVelOfCOM{0, 0, 0};
    std::vector<std::function<void(full_particle&)>> m_operations;

private:
    Derived& self() { return *static_cast<Derived*>(this); }
    virtual particleType generateParticle(size_t id) = 0;
};

// particleCount particles uniform in a ball of the given radius at the origin
class UniformSphere : public ParticleSource<Particle<POS, MASS, DENSITY>, UniformSphere> {
public:
    UniformSphere(size_t particleCount, f1_t radius, f1_t totalMass, f1_t materialDensity,
                  unsigned int seed = detail::random_seed())
        : m_radius(radius), m_matDensity(materialDensity), m_particleMass(totalMass / particleCount), m_rng(seed)
    {
        m_numberOfParticles = particleCount;
    }

private:
    Particle<POS, MASS, DENSITY> generateParticle(size_t) override
    {
        Particle<POS, MASS, DENSITY> p;
        const f1_t u1 = m_dist(m_rng);
        const f1_t u2 = m_dist(m_rng);
        const f1_t u3 = m_dist(m_rng);
        p.pos = detail::ball_point(u1, u2, u3, m_radius);
        p.mass = m_particleMass;
        p.density = m_matDensity;
        return p;
    }

    f1_t m_radius, m_matDensity, m_particleMass;
    std::default_random_engine m_rng;
    std::uniform_real_distribution<f1_t> m_dist{0.0f, 1.0f};
};

// particleCount particles of a Plummer sphere with scale radius
// plummerRadius, truncated at cutoffRadius
class PlummerSphere : public ParticleSource<Particle<POS, MASS, DENSITY>, PlummerSphere> {
public:
    PlummerSphere(size_t particleCount, f1_t plummerRadius, f1_t cutoffRadius, f1_t totalMass, f1_t materialDensity,
                  unsigned int seed = detail::random_seed())
        : m_plummerRadius(plummerRadius), m_cutoffRadius(cutoffRadius), m_matDensity(materialDensity),
          m_totalMass(totalMass), m_particleMass(totalMass / particleCount), m_rng(seed)
    {
        if (!(cutoffRadius > 0)) throw std::invalid_argument("PlummerSphere needs a cut-off radius > 0");
        m_numberOfParticles = particleCount;
    }

    // velocities of the Plummer distribution function (dynamic equilibrium)
    PlummerSphere& addRandomPlummerVelocity(f1_t G = 1, unsigned int seed = detail::random_seed())
    {
        const f1_t M = m_totalMass, a = m_plummerRadius;
        m_operations.push_back([G, M, a, rng = std::default_random_engine(seed),
                                d01 = std::uniform_real_distribution<f1_t>(0, 1),
                                d001 = std::uniform_real_distribution<f1_t>(0, 0.1f)](full_particle& p) mutable {
            const f1_t r = length(p.pos);
            f1_t x = 0, y = 0.1f;
            while (y > x * x * std::pow(1 - x * x, 3.5)) {
                x = d01(rng);
                y = d001(rng);
            }
            const f1_t u1 = d01(rng);
            const f1_t u2 = d01(rng);
            f3_t v = detail::unit_direction(u1, u2);
            v *= x * std::sqrt(2 * G * M / std::sqrt(r * r + a * a));
            p.vel += v;
        });
        return *this;
    }

private:
    Particle<POS, MASS, DENSITY> generateParticle(size_t) override
    {
        Particle<POS, MASS, DENSITY> p;
        const f1_t u1 = m_dist(m_rng);
        const f1_t u2 = m_dist(m_rng);
        p.pos = detail::unit_direction(u1, u2);
        f1_t radius = m_cutoffRadius + 1;
        while (radius > m_cutoffRadius)
            radius = m_plummerRadius / std::sqrt(std::pow(m_dist(m_rng), -2.0f / 3.0f) - 1);
        p.pos *= radius;
        p.mass = m_particleMass;
        p.density = m_matDensity;
        return p;
    }

    f1_t m_plummerRadius, m_cutoffRadius, m_matDensity, m_totalMass, m_particleMass;
    std::default_random_engine m_rng;
    std::uniform_real_distribution<f1_t> m_dist{0.0f, 1.0f};
};

}  // namespace ps

// Collects particle sources and builds one host buffer from them, in the
// order the sources were added.
template <typename hostParticleType>
class InitGenerator {
public:
    template <typename ParticleSourceType>
    void addParticles(ParticleSourceType particleSource, bool correctCOM = false, bool correctVelOfCOM = false)
    {
        static_assert(std::is_base_of<ps::detail::ParticleSourceBaseFlag, ParticleSourceType>::value,
                      "particleSource must derive from ps::ParticleSource");
        const size_t n = particleSource.getNumParticles();
        m_sources.push_back(SourceData{m_total, m_total + n, particleSource.getDesiredCOM(),
                                       particleSource.getDesiredVelOfCOM(), correctCOM, correctVelOfCOM,
                                       std::function<void(hostParticleType&, size_t, size_t)>(std::move(particleSource))});
        m_total += n;
    }

    hostParticleType generate()
    {
        hostParticleType hpb(m_total);
        hpb.initialize();
        size_t p = 0;
        for (auto& s : m_sources) {
            for (size_t j = s.start; j < s.end; ++j, ++p) s.gen(hpb, j - s.start, p);
            if (!(s.correctCOM || s.correctVelOfCOM)) continue;
            f3_t com{0, 0, 0}, vcom{0, 0, 0};
            f1_t total = 0;
            for (size_t j = s.start; j < s.end; ++j) {
                const auto q = hpb.template loadParticle<POS, MASS, VEL>(j);
                total += q.mass;
                com += q.mass * q.pos;
                vcom += q.mass * q.vel;
            }
            com /= total;
            vcom /= total;
            const f3_t dcom = s.correctCOM ? s.desiredCOM - com : f3_t{0, 0, 0};
            const f3_t dvel = s.correctVelOfCOM ? s.desiredVelOfCOM - vcom : f3_t{0, 0, 0};
            for (size_t j = s.start; j < s.end; ++j) {
                auto q = hpb.template loadParticle<POS, MASS, VEL>(j);
                q.pos += dcom;
                q.vel += dvel;
                hpb.storeParticle(j, q);
            }
        }
        return hpb;
    }

    size_t numberOfParticles() const { return m_total; }

private:
    struct SourceData {
        size_t start, end;
        f3_t desiredCOM, desiredVelOfCOM;
        bool correctCOM, correctVelOfCOM;
        std::function<void(hostParticleType&, size_t, size_t)> gen;
    };
    std::vector<SourceData> m_sources;
    size_t m_total{0};
};

#endif  // GRASPH2_INITIAL_CONDITIONS_HPP
