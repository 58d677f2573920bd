/*
 * phonon.c -- small host helpers of the phonon interpolation:
 * quadrupole_3d_to_2d (src/phonon/quadrupole_3Dto2D.c:18-117) and the
 * LO-TO direction of interpolation_driver.c:734-786.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/lelph_hip.h"

/* 3-D dynamical quadrupoles -> the 2-D ones of a slab (Appendix B of
 * Phys. Rev. X 11, 041027 (2021), Eqs. B5a, B5b, B10): with
 * Qhat_zz = (Q_zz + 2 tau_z Z_z) / eps_zz, the mixed components
 * (Q_za + tau_z Z_a) / eps_zz and (Q_az + tau_z Z_a) / eps_zz, and the
 * in-plane Q_ag + (delta_ag - eps_ag) Qhat_zz, per atom and displacement
 * direction beta.  Qpole (natom, 3, 3, 3) in place; Zborn may be NULL. */
int lelph_quadrupole_3d_to_2d(int natom, const double* epsilon, const double* atomic_pos,
                              const double* Zborn, double* Qpole)
{
    if (natom < 1 || !epsilon || !atomic_pos || !Qpole) return LELPH_ERR_ARG;
    if (!(fabs(epsilon[8]) > 0)) return LELPH_ERR_ARG;
    double* q2 = calloc(27 * (size_t)natom, sizeof(double));
    if (!q2) return LELPH_ERR_ALLOC;
    const double ie = 1.0 / epsilon[8];
    for (int k = 0; k < natom; ++k)
    {
        const double tz = atomic_pos[3 * k + 2];
        const double* Q = Qpole + 27 * k; /* Q[x][y][beta] */
        const double* Z = Zborn ? Zborn + 9 * k : NULL;
        double* O = q2 + 27 * k;
        for (int b = 0; b < 3; ++b)
        {
            const double qzz = (Q[24 + b] + 2.0 * tz * (Z ? Z[6 + b] : 0.0)) * ie;
            O[24 + b] = qzz;
            for (int a = 0; a < 2; ++a)
            {
                const double za = Z ? Z[3 * a + b] : 0.0;
                O[18 + 3 * a + b] = (Q[18 + 3 * a + b] + tz * za) * ie;
                O[9 * a + 6 + b] = (Q[9 * a + 6 + b] + tz * za) * ie;
            }
            for (int a = 0; a < 2; ++a)
                for (int g = 0; g < 2; ++g)
                    O[9 * a + 3 * g + b] =
                        Q[9 * a + 3 * g + b] + ((a == g ? 1.0 : 0.0) - epsilon[3 * a + g]) * qzz;
        }
    }
    memcpy(Qpole, q2, sizeof(double) * 27 * (size_t)natom);
    free(q2);
    return LELPH_OK;
}

/* interpolation_driver.c:734-786: Gamma approached along dir (cartesian):
 * q = 1e-5 dir / |dir| in crystal units (a_i . q), written to q_crys.
 * LELPH_ERR_ARG for |dir| < 1e-6 (the reference then switches LO-TO off). */
int lelph_loto_qpoint(const double lat_vecs[9], const double dir[3], double q_crys[3])
{
    if (!lat_vecs || !dir || !q_crys) return LELPH_ERR_ARG;
    const double n = sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
    if (n < 1e-6) return LELPH_ERR_ARG;
    double t[3];
    for (int i = 0; i < 3; ++i) t[i] = 1e-5 * dir[i] / n;
    for (int i = 0; i < 3; ++i)
        q_crys[i] = lat_vecs[i] * t[0] + lat_vecs[3 + i] * t[1] + lat_vecs[6 + i] * t[2];
    return LELPH_OK;
}
