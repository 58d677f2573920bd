/* sizeof and offsetof of every struct of include/lelph_hip.h as seen by a C
 * compiler, one "struct.field offset" line each (and "struct sizeof N"),
 * for the comparison with the ctypes mirror the Python host side uses.
 * Test code. */
#include <stddef.h>
#include <stdio.h>

#include "lelph_hip.h"

#define F(S, f) printf(#S "." #f " %zu\n", offsetof(S, f))
#define Z(S) printf(#S " sizeof %zu\n", sizeof(S))

int main(void)
{
    Z(lelph_symmetry);
    F(lelph_symmetry, Rmat);
    F(lelph_symmetry, tau);
    F(lelph_symmetry, time_rev);
    Z(lelph_lattice);
    F(lelph_lattice, natom);
    F(lelph_lattice, nsym);
    F(lelph_lattice, timerev);
    F(lelph_lattice, nspin);
    F(lelph_lattice, nspinor);
    F(lelph_lattice, nbnds);
    F(lelph_lattice, nmodes);
    F(lelph_lattice, nmag);
    F(lelph_lattice, dimension);
    F(lelph_lattice, is_soc_present);
    F(lelph_lattice, fft_dims);
    F(lelph_lattice, alat_vec);
    F(lelph_lattice, blat_vec);
    F(lelph_lattice, volume);
    F(lelph_lattice, atomic_pos);
    F(lelph_lattice, atom_type);
    F(lelph_lattice, nkpts_iBZ);
    F(lelph_lattice, kpt_iredBZ);
    F(lelph_lattice, nkpts_BZ);
    F(lelph_lattice, kpt_fullBZ_crys);
    F(lelph_lattice, kmap);
    F(lelph_lattice, syms);
    Z(lelph_local_pseudo);
    F(lelph_local_pseudo, Vloc_atomic);
    F(lelph_local_pseudo, r_grid);
    F(lelph_local_pseudo, rab_grid);
    F(lelph_local_pseudo, ngrid);
    F(lelph_local_pseudo, Zval);
    Z(lelph_pseudo);
    F(lelph_pseudo, ntype);
    F(lelph_pseudo, lmax);
    F(lelph_pseudo, nltimesj);
    F(lelph_pseudo, PP_table);
    F(lelph_pseudo, Fsign);
    F(lelph_pseudo, fCoeff);
    F(lelph_pseudo, loc_pseudo);
    F(lelph_pseudo, qmax_abs);
    Z(lelph_pair);
    F(lelph_pair, ikq);
    F(lelph_pair, ik);
    F(lelph_pair, kqsym);
    F(lelph_pair, ksym);
    Z(lelph_tuning);
    F(lelph_tuning, split_target);
    F(lelph_tuning, fft_lines);
    F(lelph_tuning, fp32_form);
    F(lelph_tuning, ctile);
    F(lelph_tuning, nl_form);
    Z(lelph_longrange);
    F(lelph_longrange, epsilon);
    F(lelph_longrange, Zborn);
    F(lelph_longrange, Qpole);
    F(lelph_longrange, Zvals);
    F(lelph_longrange, only_induced);
    F(lelph_longrange, eta_bare);
    F(lelph_longrange, eta_induced);
    F(lelph_longrange, EcutRy);
    F(lelph_longrange, nmags_add);
    Z(lelph_qexml_info);
    F(lelph_qexml_info, natom);
    F(lelph_qexml_info, ntype);
    F(lelph_qexml_info, nsym);
    F(lelph_qexml_info, fft_dims);
    F(lelph_qexml_info, lat_vec);
    F(lelph_qexml_info, alat);
    F(lelph_qexml_info, dimension);
    F(lelph_qexml_info, is_soc_present);
    F(lelph_qexml_info, nmag);
    F(lelph_qexml_info, nspinor);
    F(lelph_qexml_info, ph_mag_symm);
    F(lelph_qexml_info, ph_tim_rev);
    F(lelph_qexml_info, pseudo_dir);
    return 0;
}
