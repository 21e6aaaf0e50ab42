// Particle sources sampled on the device (start_generate): the reference's
// point and surface sources (ParticleGeneratorHost.cpp:33-159,
// ParticleGeneratorDevice.cu:56-304), its start-point displacement
// (ParticleDisplacer.cpp:20-31, 48-134) and the energy / direction samplers
// the workloads need, drawn straight into the context's FP64 arrays.  Their
// sin, cos, acos, atan2, log and exp are the shared deterministic ones of
// detmath.h, which the oracle's restatement calls too.
#include "kernels.h"
#include "sources.h"

namespace sg {

namespace {

constexpr double kEvJ = 1.60217733000001e-19;  // Constants.hpp:15 (eV_J)

// Two 53-bit uniforms in [0,1): counter {particle, 1, source, draw}, key
// {seed_lo, seed_hi}.  The 1 keeps them apart from the step's variates, whose
// second counter word is 0.
__device__ inline void src_uniforms(uint64_t seed, uint32_t i, uint32_t s, uint32_t draw, double& ua, double& ub)
{
    uint32_t w[4];
    philox4x32_10(i, 1u, s, draw, (uint32_t)seed, (uint32_t)(seed >> 32), w);
    const double two_m53 = 1.0 / 9007199254740992.0;
    ua = (double)(((uint64_t)w[0] << 21) | (w[1] >> 11)) * two_m53;
    ub = (double)(((uint64_t)w[2] << 21) | (w[3] >> 11)) * two_m53;
}

// standard normal from two uniforms (Box-Muller, cosine branch)
__device__ inline double normal_bm(double ua, double ub)
{
    return sqrt(-2.0 * dm_log(1.0 - ua)) * dm_cos(2.0 * kPi * ub);
}

__global__ void generate_kernel(SourceK S, uint32_t s_index, uint64_t seed, const double* __restrict__ centres,
                                const double* __restrict__ normals, double* __restrict__ x, double* __restrict__ y,
                                double* __restrict__ z, double* __restrict__ vx, double* __restrict__ vy,
                                double* __restrict__ vz, int32_t* __restrict__ flags)
{
    const int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (j >= S.count) return;
    const int64_t i = S.first + j;
    double u0, u1, u2, u3;
    src_uniforms(seed, (uint32_t)i, s_index, 0u, u0, u1);
    src_uniforms(seed, (uint32_t)i, s_index, 1u, u2, u3);

    // position and source direction
    double px, py, pz, t0, p0, dx, dy, dz;
    if (S.kind == START_SOURCE_POINT) {
        px = S.pos[0];
        py = S.pos[1];
        pz = S.pos[2];
        t0 = S.theta;
        p0 = S.phi;
        // ParticleDisplacer.cpp:68-71
        const double st = dm_sin(t0);
        dx = st * dm_cos(p0);
        dy = st * dm_sin(p0);
        dz = dm_cos(t0);
    } else {
        // count / n_cells particles per cell, one more for the first count % n_cells
        // cells of the (host-shuffled) cell order
        const int64_t q = S.count / S.n_cells, r = S.count % S.n_cells;
        const int64_t cell = j < (q + 1) * r ? j / (q + 1) : r + (j - (q + 1) * r) / q;
        const double* cc = centres + 3 * (S.cell0 + cell);
        const double* nn = normals + 3 * (S.cell0 + cell);
        px = cc[0];
        py = cc[1];
        pz = cc[2];
        // ParticleGeneratorHost.cpp:130-146: normalised normal, theta = acos(n_z), phi = atan2(n_y, n_x)
        const double mag = sqrt(nn[0] * nn[0] + nn[1] * nn[1] + nn[2] * nn[2]);
        const double nx = nn[0] / mag, ny = nn[1] / mag, nz = nn[2] / mag;
        t0 = dm_acos(nz);
        p0 = dm_atan2(ny, nx);
        dx = nn[0];  // the displacer takes the normal as given (ParticleDisplacer.cpp:110-113)
        dy = nn[1];
        dz = nn[2];
    }
    if (S.displace != 0.0) {  // ParticleDisplacer.cpp:20-31
        const double mc = fmax(fmax(fabs(dx), fabs(dy)), fabs(dz));
        if (!(mc < 2.220446049250313e-16)) {
            px = px + (dx / mc) * S.displace;
            py = py + (dy / mc) * S.displace;
            pz = pz + (dz / mc) * S.displace;
        }
    }

    double wx, wy, wz;
    if (S.energy == START_ENERGY_MAXWELLIAN) {
        // each component normal with variance kT/m
        const double sig = sqrt((S.p0 * kEvJ) / S.mass);
        wx = sig * normal_bm(u0, u1);
        wy = sig * (sqrt(-2.0 * dm_log(1.0 - u0)) * dm_sin(2.0 * kPi * u1));
        wz = sig * normal_bm(u2, u3);
    } else {
        double e;  // eV
        if (S.energy == START_ENERGY_MONO) {
            e = S.energy_eV;
        } else if (S.energy == START_ENERGY_THOMPSON) {
            // F(E) = E^2 / (E + U)^2 on [0, E_max]: E = U r / (1 - r), r = sqrt(u F(E_max))
            const double fr = S.p1 / (S.p1 + S.p0);
            const double rr = sqrt(u0 * (fr * fr));
            e = S.p0 * rr / (1.0 - rr);
        } else {  // lognormal with mean p0 and log-standard deviation p1
            const double mu = dm_log(S.p0) - 0.5 * S.p1 * S.p1;
            e = dm_exp(mu + S.p1 * normal_bm(u0, u1));
        }
        // ParticleDynamicUtils.hpp:52-96: E in J, v = sqrt(2 E / m)
        const double ej = e * kEvJ;
        const double v = sqrt(2 * ej / S.mass);
        if (S.angle == START_ANGLE_FAN) {
            // polar angle theta + u * expansion (u uniform in [-1, 1)), azimuth phi
            const double th = S.expansion != 0.0 ? t0 + (-1.0 + 2.0 * u2) * S.expansion : t0;
            wx = v * dm_sin(th) * dm_cos(p0);
            wy = v * dm_sin(th) * dm_sin(p0);
            wz = v * dm_cos(th);
        } else {
            double ct, ph;
            if (S.angle == START_ANGLE_COSINE) {
                ct = sqrt(u2);
            } else if (S.angle == START_ANGLE_ISOTROPIC) {
                ct = 1.0 - 2.0 * u2;
            } else {  // cone of half-angle expansion
                ct = 1.0 - u2 * (1.0 - dm_cos(S.expansion));
            }
            ph = 2.0 * kPi * u3;
            const double st = sqrt(fmax(0.0, 1.0 - ct * ct));
            const double a = st * dm_cos(ph), b = st * dm_sin(ph);
            if (S.angle == START_ANGLE_ISOTROPIC) {
                wx = v * a;
                wy = v * b;
                wz = v * ct;
            } else {
                // local (a, b, ct) about d = (sin t0 cos p0, sin t0 sin p0, cos t0):
                // e1 = (cos t0 cos p0, cos t0 sin p0, -sin t0), e2 = (-sin p0, cos p0, 0)
                const double s0 = dm_sin(t0), c0 = dm_cos(t0), sp = dm_sin(p0), cp = dm_cos(p0);
                wx = v * ((a * (c0 * cp) - b * sp) + ct * (s0 * cp));
                wy = v * ((a * (c0 * sp) + b * cp) + ct * (s0 * sp));
                wz = v * (ct * c0 - a * s0);
            }
        }
    }
    x[i] = px;
    y[i] = py;
    z[i] = pz;
    vx[i] = wx;
    vy[i] = wy;
    vz[i] = wz;
    flags[i] = S.flags;
}

inline unsigned blocks(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

}  // namespace

void launch_generate(hipStream_t s, const SourceK& S, uint32_t s_index, uint64_t seed, const double* centres,
                     const double* normals, double* x, double* y, double* z, double* vx, double* vy, double* vz,
                     int32_t* flags)
{
    if (S.count > 0)
        generate_kernel<<<blocks(S.count, 256), 256, 0, s>>>(S, s_index, seed, centres, normals, x, y, z, vx, vy,
                                                              vz, flags);
}

}  // namespace sg
