/* Prints, as JSON, the sizes of the hot-path structs of sepstrct.h and the
 * offsets of their fields, for whichever include directory it is compiled
 * against (-I).  Used to pin this repository's header to the reference's. */
#include <stddef.h>
#include <stdio.h>

#include "sepstrct.h"

#define F(T, f) printf("%s\"%s.%s\": %lu", first ? "" : ",\n", #T, #f, (unsigned long)offsetof(T, f)), first = 0
#define Z(T) printf("%s\"sizeof.%s\": %lu", first ? "" : ",\n", #T, (unsigned long)sizeof(T)), first = 0

int main(void)
{
	int first = 1;
	printf("{\n");
	Z(seppart);
	F(seppart, x); F(seppart, v); F(seppart, f); F(seppart, a); F(seppart, m);
	F(seppart, type); F(seppart, z); F(seppart, neighb); F(seppart, cross_neighb);
	F(seppart, crossings); F(seppart, molindex); F(seppart, bond); F(seppart, angle);
	F(seppart, dihed); F(seppart, xn); F(seppart, prevf);
	Z(sepsys);
	F(sepsys, npart); F(sepsys, length); F(sepsys, volume); F(sepsys, dt); F(sepsys, tnow);
	F(sepsys, ndof); F(sepsys, max_dist2); F(sepsys, cf); F(sepsys, lsubbox);
	F(sepsys, nsubbox); F(sepsys, skin); F(sepsys, neighb_update); F(sepsys, neighb_flag);
	F(sepsys, nupdate_neighb); F(sepsys, omp_flag); F(sepsys, molptr);
	Z(sepret);
	F(sepret, ekin); F(sepret, epot); F(sepret, ecoul); F(sepret, P); F(sepret, kin_P);
	F(sepret, pot_P); F(sepret, p); F(sepret, pot_P_bond); F(sepret, t_mol);
	Z(sepmolinfo);
	F(sepmolinfo, num_mols); F(sepmolinfo, num_bonds); F(sepmolinfo, blist);
	F(sepmolinfo, num_angles); F(sepmolinfo, alist); F(sepmolinfo, num_dihedrals);
	F(sepmolinfo, dlist); F(sepmolinfo, blengths); F(sepmolinfo, angles);
	F(sepmolinfo, dihedrals);
	printf("\n}\n");
	return 0;
}
