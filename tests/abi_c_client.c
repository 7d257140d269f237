/* A plain C11 client of include/masx.h: builds a 1-D chain of 3x3 blocks
 * (tridiagonal, 100 block rows), analyses it with the MAS preconditioner over
 * an identity partition and prints the plan sizes and the hierarchy's level
 * counts.  Needs no GPU (masx_analyze is host-only).  Exit status 0 on
 * success. */
#include <stdio.h>
#include <stdlib.h>

#include "masx.h"

#define N 100

static int fail(const char* what)
{
    fprintf(stderr, "%s: %s\n", what, masx_last_error());
    return 1;
}

int main(void)
{
    int      rows[2 * N], cols[2 * N];
    int      nb = 0;
    uint32_t start[N], num[N], list[2 * N];
    int      n_list = 0;
    int      part_to_real[128], real_to_part[N];
    for(int i = 0; i < N; ++i)
    {
        rows[nb] = i, cols[nb] = i, ++nb;
        if(i + 1 < N) rows[nb] = i, cols[nb] = i + 1, ++nb;
        start[i] = (uint32_t)n_list;
        if(i > 0) list[n_list++] = (uint32_t)(i - 1);
        if(i + 1 < N) list[n_list++] = (uint32_t)(i + 1);
        num[i] = (uint32_t)n_list - start[i];
        real_to_part[i] = i;
    }
    for(int k = 0; k < 128; ++k) part_to_real[k] = k < N ? k : -1;

    masx_solver* s = NULL;
    if(masx_create(&s)) return fail("masx_create");
    masx_system_desc d = {N, nb, 0, N, 1, 0, 32, 0};
    int sys = -1;
    if(masx_add_system(s, &d, rows, cols, &sys)) return fail("masx_add_system");
    if(masx_set_topology(s, sys, start, num, list, n_list, part_to_real, 128, real_to_part))
        return fail("masx_set_topology");
    if(masx_analyze(s)) return fail("masx_analyze");
    masx_plan_info info;
    if(masx_plan_info_get(s, &info)) return fail("masx_plan_info_get");
    int levels = 0, total = 0, level_begin[8] = {0};
    if(masx_get_hierarchy(s, sys, &levels, &total, NULL, level_begin, NULL, NULL, NULL, NULL))
        return fail("masx_get_hierarchy");
    int level_count[8] = {0};
    if(masx_get_hierarchy(s, sys, &levels, &total, level_count, level_begin, NULL, NULL, NULL, NULL))
        return fail("masx_get_hierarchy");
    printf("rows %d nnz %lld nb %lld banks %d coarse %d levels %d total %d counts", info.n_rows, info.nnz,
           info.nb_total, info.n_banks, info.n_coarse, levels, total);
    for(int l = 1; l <= levels; ++l) printf(" %d", level_count[l]);
    printf("\n");
    /* a bad call reports an error instead of aborting */
    if(masx_add_system(s, NULL, rows, cols, &sys) == 0) return 1;
    if(masx_last_error()[0] == '\0') return 1;
    masx_destroy(s);
    return 0;
}
