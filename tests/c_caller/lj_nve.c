/*
 * A C99 program against include/ and libsep.so, in the call order of the
 * reference's own CPU driver (cuda/tcpu_0.c:27-60): sep_init_xyz,
 * sep_sys_setup, sep_set_omp, then per step sep_reset_retval,
 * sep_reset_force, sep_force_pairs(..., sep_lj_shift, ..., SEP_ALL) and
 * sep_leapfrog, and at the end sep_save_xyz, sep_close and sep_free_sys.
 * The only addition is sep_set_skin, to run the skin of the parity
 * fixture it is compared with.
 *
 *     lj_nve <in.xyz> <steps> <skin> <out.xyz>
 *
 * Prints, per step, "step n epot ekin pot_P[9] kin_P[9] tnow" with every
 * double as %.17g, and exits with the library's error code (0 on success).
 */
#include "sep.h"

int main(int argc, char **argv)
{
	sepatom *atoms;
	sepsys sys;
	sepret ret;
	double lbox[3];
	int natoms;

	if (argc != 5) {
		fprintf(stderr, "usage: %s in.xyz steps skin out.xyz\n", argv[0]);
		return 2;
	}
	const int nloops = atoi(argv[2]);
	const double rcut = 2.5, dt = 0.005;

	atoms = sep_init_xyz(lbox, &natoms, argv[1], 'v');
	sys = sep_sys_setup(lbox[0], lbox[1], lbox[2], rcut, dt, natoms, SEP_LLIST_NEIGHBLIST);
	sep_set_omp(4, &sys);
	sep_set_skin(&sys, atof(argv[3]));

	for (int n = 0; n < nloops; n++) {
		sep_reset_retval(&ret);
		sep_reset_force(atoms, &sys);
		sep_force_pairs(atoms, "AA", rcut, sep_lj_shift, &sys, &ret, SEP_ALL);
		sep_leapfrog(atoms, &sys, &ret);
		printf("step %d %.17g %.17g", n, ret.epot, ret.ekin);
		for (int a = 0; a < 3; a++)
			for (int b = 0; b < 3; b++)
				printf(" %.17g", ret.pot_P[a][b]);
		for (int a = 0; a < 3; a++)
			for (int b = 0; b < 3; b++)
				printf(" %.17g", ret.kin_P[a][b]);
		printf(" %.17g\n", sys.tnow);
	}

	sep_save_xyz(atoms, "A", argv[4], "w", &sys);
	const int err = sep_gpu_last_error();
	sep_close(atoms, natoms);
	sep_free_sys(&sys);
	return err;
}
