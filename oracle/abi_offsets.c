/* Prints the byte offset of every field of the two ABI structs. Built twice by
 * oracle/Makefile: against include/libcusna.h and against the reference
 * header, so a test can show the layouts are identical. Test infrastructure. */
#include <stddef.h>
#include <stdio.h>
#include LIBCUSNA_HEADER

#define P(T, f) printf("%s.%s %zu\n", #T, #f, offsetof(T, f))

int main(void) {
  printf("sizeof.problem %zu\nsizeof.result %zu\n", sizeof(cusna_saom_problem),
         sizeof(cusna_saom_result));
  P(cusna_saom_problem, struct_version); P(cusna_saom_problem, n); P(cusna_saom_problem, W);
  P(cusna_saom_problem, k); P(cusna_saom_problem, M); P(cusna_saom_problem, threads);
  P(cusna_saom_problem, x0_out); P(cusna_saom_problem, x0_in); P(cusna_saom_problem, sim_out);
  P(cusna_saom_problem, sim_in); P(cusna_saom_problem, lam); P(cusna_saom_problem, theta);
  P(cusna_saom_problem, type_ids); P(cusna_saom_problem, etypes); P(cusna_saom_problem, wtab);
  P(cusna_saom_problem, p1); P(cusna_saom_problem, p2); P(cusna_saom_problem, vmat);
  P(cusna_saom_problem, vmiss); P(cusna_saom_problem, dyc_idx); P(cusna_saom_problem, dyc);
  P(cusna_saom_problem, int0); P(cusna_saom_problem, int1); P(cusna_saom_problem, int2);
  P(cusna_saom_problem, dyn_flags); P(cusna_saom_problem, kdist); P(cusna_saom_problem, keep_out);
  P(cusna_saom_problem, keep_in); P(cusna_saom_problem, force_out); P(cusna_saom_problem, force_in);
  P(cusna_saom_problem, perm); P(cusna_saom_problem, od0); P(cusna_saom_problem, id0);
  P(cusna_saom_problem, maxdegree); P(cusna_saom_problem, need_mask);
  P(cusna_saom_problem, conditional); P(cusna_saom_problem, target_dist);
  P(cusna_saom_problem, symflag); P(cusna_saom_problem, seed); P(cusna_saom_problem, has_comp);
  P(cusna_saom_problem, nact); P(cusna_saom_problem, act); P(cusna_saom_problem, inactw);
  P(cusna_saom_problem, kb); P(cusna_saom_problem, lamb); P(cusna_saom_problem, thetab);
  P(cusna_saom_problem, bty); P(cusna_saom_problem, z0); P(cusna_saom_problem, vb);
  P(cusna_saom_problem, bkdist); P(cusna_saom_problem, zmin); P(cusna_saom_problem, zmax);
  P(cusna_saom_problem, zbar); P(cusna_saom_problem, brange); P(cusna_saom_problem, bsimmean);
  P(cusna_saom_problem, nrate); P(cusna_saom_problem, rkind); P(cusna_saom_problem, ralpha);
  P(cusna_saom_problem, ru); P(cusna_saom_problem, vmat64); P(cusna_saom_problem, wtab64);
  P(cusna_saom_problem, vb64); P(cusna_saom_problem, p1_64); P(cusna_saom_problem, p2_64);
  P(cusna_saom_result, struct_version); P(cusna_saom_result, stats); P(cusna_saom_result, ntim);
  P(cusna_saom_result, scores); P(cusna_saom_result, dist); P(cusna_saom_result, dist_stat);
  P(cusna_saom_result, steps); P(cusna_saom_result, bdist); P(cusna_saom_result, bscores);
  P(cusna_saom_result, z_end); P(cusna_saom_result, rscores); P(cusna_saom_result, actor_dist);
  P(cusna_saom_result, want_end_nets); P(cusna_saom_result, a_ends);
  printf("sizeof.multi_problem %zu\nsizeof.multi_result %zu\n",
         sizeof(cusna_saom_multi_problem), sizeof(cusna_saom_multi_result));
  P(cusna_saom_multi_problem, struct_version); P(cusna_saom_multi_problem, n);
  P(cusna_saom_multi_problem, W); P(cusna_saom_multi_problem, k); P(cusna_saom_multi_problem, M);
  P(cusna_saom_multi_problem, D); P(cusna_saom_multi_problem, threads);
  P(cusna_saom_multi_problem, x0_out); P(cusna_saom_multi_problem, x0_in);
  P(cusna_saom_multi_problem, lam); P(cusna_saom_multi_problem, theta);
  P(cusna_saom_multi_problem, type_ids); P(cusna_saom_multi_problem, etypes);
  P(cusna_saom_multi_problem, netof); P(cusna_saom_multi_problem, netref);
  P(cusna_saom_multi_problem, wtab); P(cusna_saom_multi_problem, p1);
  P(cusna_saom_multi_problem, p2); P(cusna_saom_multi_problem, vmat);
  P(cusna_saom_multi_problem, dyc); P(cusna_saom_multi_problem, dyc_idx);
  P(cusna_saom_multi_problem, perm); P(cusna_saom_multi_problem, symf);
  P(cusna_saom_multi_problem, seed); P(cusna_saom_multi_problem, act);
  P(cusna_saom_multi_problem, nact);
  P(cusna_saom_multi_result, struct_version); P(cusna_saom_multi_result, rscores);
  P(cusna_saom_multi_result, scores); P(cusna_saom_multi_result, ntim);
  P(cusna_saom_multi_result, steps); P(cusna_saom_multi_result, a_ends);
  return 0;
}
