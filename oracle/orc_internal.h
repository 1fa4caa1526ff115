/* ORACLE (test infrastructure only): internal structures shared by the
 * oracle's C files. */
#ifndef ORC_INTERNAL_H
#define ORC_INTERNAL_H

#define ORC_MAXPRIM 16
#define ORC_MAXL 4

typedef struct {
    int L;
    double center[3];
    int nprim;
    double exps[ORC_MAXPRIM];
    double c_radial[ORC_MAXPRIM];
    int ncomp;
    int comp[15];         /* basis (Cartesian AO) index of each component */
    int ang[15][3];
    double n_ang[15];
    int sph_base;
    double c2s[9 * 15];      /* [2L+1][ncart] in local descending-lex order */
    double c2s_comp[9 * 15]; /* [2L+1][ncomp] in component order */
} orc_group;

struct orc_basis {
    int ncart;
    int nsph;
    int ngroups;
    orc_group* groups;
};

void orc_set_error(const char* msg);
double orc_dfact(int n);
int orc_ncart_L(int L);
int orc_cart_local(int L, int lx, int ly);
void orc_cart_tuple(int L, int idx, int* t);
void orc_rys_init(void);

#endif
