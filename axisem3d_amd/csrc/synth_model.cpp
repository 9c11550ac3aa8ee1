// synth_model.cpp -- host generator of the synthetic 3D isotropic model
// (preloop side, OpenMP).  Computes, for a range of 3D elements, the
// per-slice lambda, mu (with the unrelaxed-modulus correction of the
// attenuation) and the memory-variable moduli dlambda, dmu, exactly as
// axisem3d_amd.preloop.make_3d_isotropic does in numpy:
//   d(x) = amp sqrt(2/K) sum_k cos(w_k . x + b_k), clipped to +-clip,
//   at x = (s cos phi_j, s sin phi_j, z), phi_j = 2 pi j / Nr
//   (Model3D.cpp:32,87); lambda, mu scaled by (1 + d)^2;
//   attenuation: Material::createIsotropic + AttBuilder::createAttenuation
//   (src/preloop/se_model/quad/physical/Material.cpp:585-598,
//   AttBuilder.cpp:103-178).
#include <cmath>
#include <cstdint>
#include <vector>

extern "C" {

int axs3d_synth_iso3d(int ne, const int32_t *nr, const double *sz,
                      const double *lam_mu1d, const double *vp1d, int nfeat,
                      const double *feat, double amp, double clip,
                      int att_mode, const double *qkp, const double *qmu,
                      const double *wcg4, double ysum, double w10f,
                      double fact, double *prop_out, double *att_out,
                      double *vmax_out) {
    const int N = 25;
    const int cg4pt[4] = {6, 8, 16, 18};
    const int npa = att_mode == 1 ? 4 : N;
    std::vector<int64_t> off(ne + 1, 0);
    for (int e = 0; e < ne; e++) off[e + 1] = off[e] + nr[e];
    const double norm = amp * std::sqrt(2.0 / nfeat);
#pragma omp parallel for schedule(dynamic, 16)
    for (int e = 0; e < ne; e++) {
        const int n = nr[e];
        const double *s = sz + (int64_t)e * 2 * N;
        const double *z = s + N;
        const double *lam1 = lam_mu1d + (int64_t)e * 2 * N;
        const double *mu1 = lam1 + N;
        double vmax = 0.;
        for (int j = 0; j < n; j++) {
            const double phi = 2.0 * M_PI * j / n;
            const double c = std::cos(phi), sn = std::sin(phi);
            double lam[25], mu[25];
            for (int p = 0; p < N; p++) {
                const double x = s[p] * c, y = s[p] * sn;
                double acc = 0.;
                for (int k = 0; k < nfeat; k++) {
                    const double *f = feat + 4 * k;
                    acc += std::cos(x * f[0] + y * f[1] + z[p] * f[2] + f[3]);
                }
                double d = norm * acc;
                d = d < -clip ? -clip : (d > clip ? clip : d);
                const double sc = (1.0 + d) * (1.0 + d);
                lam[p] = lam1[p] * sc;
                mu[p] = mu1[p] * sc;
                const double v = vp1d[(int64_t)e * N + p] * std::sqrt(sc);
                vmax = v > vmax ? v : vmax;
            }
            double *po = prop_out + (off[e] + j) * 2 * N;
            if (att_mode < 0) {
                for (int p = 0; p < N; p++) {
                    po[p] = lam[p];
                    po[N + p] = mu[p];
                }
                continue;
            }
            double kp[25], dlm[25], dmu[25], kp_na[25], kp_at[25], mu_na[25],
                mu_at[25];
            for (int p = 0; p < N; p++) {
                const double qk = qkp[(int64_t)e * N + p];
                const double qm = qmu[(int64_t)e * N + p];
                kp[p] = lam[p] + (2.0 / 3.0) * mu[p];
                kp_na[p] = 1.0 + w10f / qk;
                double dk = kp_na[p] / (qk / ysum + (1.0 - fact));
                kp_at[p] = kp_na[p] + dk * fact;
                dk = dk * kp[p];
                mu_na[p] = 1.0 + w10f / qm;
                double dm = mu_na[p] / (qm / ysum + (1.0 - fact));
                mu_at[p] = mu_na[p] + dm * fact;
                dm = dm * mu[p];
                dlm[p] = dk - (2.0 / 3.0) * dm;
                dmu[p] = dm;
            }
            double *ao = att_out + (off[e] + j) * 2 * npa;
            if (att_mode == 1) {
                const double *w4 = wcg4 + (int64_t)e * 4;
                for (int q = 0; q < 4; q++) {
                    ao[q] = w4[q] * dlm[cg4pt[q]];
                    ao[4 + q] = w4[q] * dmu[cg4pt[q]];
                }
                for (int p = 0; p < N; p++) {
                    kp[p] = kp[p] * kp_na[p];
                    mu[p] = mu[p] * mu_na[p];
                }
                for (int q = 0; q < 4; q++) {
                    const int p = cg4pt[q];
                    kp[p] *= 1.0 + w4[q] * (kp_at[p] / kp_na[p] - 1.0);
                    mu[p] *= 1.0 + w4[q] * (mu_at[p] / mu_na[p] - 1.0);
                }
            } else {
                for (int p = 0; p < N; p++) {
                    ao[p] = dlm[p];
                    ao[N + p] = dmu[p];
                    kp[p] = kp[p] * kp_at[p];
                    mu[p] = mu[p] * mu_at[p];
                }
            }
            for (int p = 0; p < N; p++) {
                po[p] = kp[p] - (2.0 / 3.0) * mu[p];
                po[N + p] = mu[p];
            }
        }
        if (vmax_out) vmax_out[e] = vmax;
    }
    return 0;
}

}  // extern "C"
