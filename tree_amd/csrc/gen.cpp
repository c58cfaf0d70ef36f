// gen.cpp -- seeded synthetic plasma (host side of libtreemi.so).
//
// Positions: a spherical Gaussian per species, normals from Philox4x32-10
// (Salmon et al., "Parallel random numbers: as easy as 1, 2, 3", SC'11) with
// the Box-Muller transform.  Velocities follow Vinit (main.f90:709-767):
// Gaussian draws, then one rescale per species so that sum m v^2 equals
// k_B (3N - Nref) T exactly (Nref = 6, data.f90:38; k_B = 8.617343e-5 eV/K,
// main.f90:733).
#include <cmath>
#include <cstdint>
#include <vector>

#include "../../include/treemi.h"

namespace {

struct P4 {
    uint32_t v[4];
};

// Philox4x32 with 10 rounds: counter c, key k
P4 philox4x32_10(P4 c, uint32_t k0, uint32_t k1)
{
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
    for (int r = 0; r < 10; r++) {
        if (r > 0) {
            k0 += W0;
            k1 += W1;
        }
        uint64_t p0 = (uint64_t)M0 * c.v[0];
        uint64_t p1 = (uint64_t)M1 * c.v[2];
        P4 o;
        o.v[0] = (uint32_t)(p1 >> 32) ^ c.v[1] ^ k0;
        o.v[1] = (uint32_t)p1;
        o.v[2] = (uint32_t)(p0 >> 32) ^ c.v[3] ^ k1;
        o.v[3] = (uint32_t)p0;
        c = o;
    }
    return c;
}

inline double u53(uint32_t hi, uint32_t lo)
{
    return (double)((((uint64_t)hi << 32) | lo) >> 11) * (1.0 / 9007199254740992.0);
}

// six standard normals for (particle i, species s): 3 for position, 3 for velocity
void normals6(uint64_t seed, uint32_t i, uint32_t s, double out[6])
{
    for (uint32_t j = 0; j < 3; j++) {
        P4 c = {{i, s, j, 0x5ca1ab1eu}};
        P4 r = philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
        double u1 = 1.0 - u53(r.v[0], r.v[1]);  // (0, 1]
        double u2 = u53(r.v[2], r.v[3]);
        double rad = std::sqrt(-2.0 * std::log(u1));
        double ang = 6.283185307179586476925286766559 * u2;
        out[2 * j] = rad * std::cos(ang);
        out[2 * j + 1] = rad * std::sin(ang);
    }
}

}  // namespace

extern "C" {

// exposed for the generator's known-answer test
void tb_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    P4 c = {{ctr[0], ctr[1], ctr[2], ctr[3]}};
    P4 r = philox4x32_10(c, key[0], key[1]);
    for (int k = 0; k < 4; k++) out[k] = r.v[k];
}

double tb_kelbg_length(double te_kelvin, double xm_e, double xm_i)
{
    const double hbar = 1.054571817e-34;    // J s
    const double amu = 1.66053906660e-27;   // kg
    const double kb = 1.380649e-23;         // J/K
    const double mu = xm_e * xm_i / (xm_e + xm_i) * amu;
    return hbar / std::sqrt(2.0 * mu * kb * te_kelvin) * 1.0e6;  // um
}

int tb_gen_plasma(uint64_t seed, int32_t ne, int32_t ni, double sigma, double te, double ti, double xm_e,
                  double xm_i, tb_pt *e, tb_pt *i)
{
    if (ne < 1 || ni < 1 || !e || !i || !(sigma > 0.0)) return TB_ERR_ARG;
    const double kb = 8.617343e-5;  // eV/K, main.f90:733
    const int32_t n[2] = {ne, ni};
    tb_pt *dst[2] = {e, i};
    const double T[2] = {te, ti}, xm[2] = {xm_e, xm_i};
    for (int s = 0; s < 2; s++) {
        double mv2 = 0.0;
        for (int32_t j = 0; j < n[s]; j++) {
            double g[6];
            normals6(seed, (uint32_t)j, (uint32_t)s, g);
            tb_pt &p = dst[s][j];
            for (int k = 0; k < 3; k++) {
                p.xx[k] = sigma * g[k];
                p.xv[k] = g[3 + k] * std::sqrt(T[s]);  // fluct(SQRT(T)), main.f90:727
                p.ff[k] = 0.0;
                mv2 = mv2 + xm[s] * p.xv[k] * p.xv[k];
            }
            p.id = j + 1;
            p.pad_ = 0;
        }
        // exact-temperature rescale, main.f90:731-737
        double dof = (double)(3 * (int64_t)n[s] - 6);
        if (dof > 0.0 && T[s] > 0.0) {
            double T0 = mv2 / kb / dof;
            double lam = std::sqrt(T[s] / T0);
            for (int32_t j = 0; j < n[s]; j++)
                for (int k = 0; k < 3; k++) dst[s][j].xv[k] = dst[s][j].xv[k] * lam;
        }
    }
    return TB_OK;
}

}  // extern "C"
