// kern_sw.hip -- panel-pair integrals in the short-wavelength regime
// (|k| R_panel > pi for the pair's larger panel), every frequency.
//
// GetPanelPanelInteractions (PanelPanelInteractions.cc:426-491) takes a
// near pair out of the desingularization branch when |k| max(Ra, Rb) > pi:
//   * no common vertex: the 78-point x 78-point rule without
//     desingularization (GetPPIs_Cubature with HighOrder);
//   * common edge or vertex: the generalized Taylor-Duffy method with
//     P = {1, PMCHWG1, PMCHWC} and the Helmholtz / gradient-Helmholtz
//     kernels, H0 = I1 - 4 I0/k^2, H1 = I2 (up to |k| R = 20; past it the
//     reference's high-k gradient kernel reads outside its factorial
//     table, TaylorDuffy.cc:777,812, and the host rejects the pair).
// The Taylor-Duffy integrand follows TaylorDuffy.cc:152-258, :373-935,
// :940-1262, :1265-1528; its p-adaptive Clenshaw-Curtis cubature (up to
// three dimensions, relative tolerance 1e-10, at most 1000 new points)
// follows libs/libSGJC/pcubature.c:60-430, visiting cached values in the
// same order.  One wave per (pair, region, Qa, Qb) job.
#include <hip/hip_runtime.h>

#include "bem.h"
#include "tcr20.inc"

namespace scuffx {

#define SCUFFX_INV4PI 0.079577471545947667884   // 1/(4 pi)
#define TD_CT 3
#define TD_CE 2
#define TD_CV 1
#define TD_NREG 6
#define TD_NWP 5
#define TD_NMON 22
#define TD_NPK 3
#define TD_MAXEVAL 1000
#define TD_RELTOL 1.0e-10
#define PC_M 10                          // highest rule order a job can reach
#define PC_NX (1 << (PC_M - 1))
#define PC_NW (PC_M + (1 << PC_M) + (1 << PC_M) + 1)
#define PC_MAXVAL 2112                   // cached points per job (bound: 27 + 1000 + 1027)
#define PC_MAXENT 48                     // cache entries per job
#define SW_FDIM (2 * TD_NPK)
// scratch slab of one job (its cached integrand values), in doubles
#define SW_SLAB (PC_MAXVAL * SW_FDIM)

enum { T1 = 0, TY1 = 1, TY12 = 2, TY2 = 3, TY1Y2 = 4, TY22 = 5, TY13 = 6,
       TY12Y2 = 7, TY1Y22 = 8, TY23 = 9, TY12Y22 = 10, TY13Y2 = 11, TY13Y22 = 12,
       TY3 = 13, TY1Y3 = 14, TY2Y3 = 15, TY12Y3 = 16, TY1Y2Y3 = 17, TY22Y3 = 18,
       TY12Y2Y3 = 19, TY1Y22Y32 = 20, TY22Y32 = 21 };
enum { K_HELM = 0, K_GRADHELM = 1 };

__constant__ double c_swx[PC_NX];
__constant__ double c_sww[PC_NW];
__constant__ double c_swtcr20[234];

struct TDGeo {
  int WhichCase, NumPKs;
  int KIndex[TD_NPK];
  cplx k;
  double A2, B2, AP2, BP2, L2, AdB, AdAP, AdBP, AdL, BdAP, BdBP, APdBP, BPdL;
  int nMin[TD_NPK], nMax[TD_NPK];
};

// Upsilon coefficient tables (TaylorDuffy.cc:940-1262); C = {C2010, C2001,
// C0210, C0201}, M symmetric 4x4 (upper triangle used, M[i + 4 j]), V[4], S
__device__ void td_upsilon(int WhichCase, const double *C, const double *M, const double *V,
                           double S, double (*U)[TD_NWP][TD_NMON])
{
  for (int d = 0; d < TD_NREG; d++)
    for (int n = 0; n < TD_NWP; n++)
      for (int q = 0; q < TD_NMON; q++) U[d][n][q] = 0.0;
  double C2010 = C[0], C2001 = C[1], C0210 = C[2], C0201 = C[3];
  double M11 = M[0], M12 = M[4], M13 = M[8], M14 = M[12];
  double M22 = M[5], M23 = M[9], M24 = M[13];
  double M33 = M[10], M34 = M[14];
  double M44 = M[15];
  double V1 = V[0], V2 = V[1], V3 = V[2], V4 = V[3];

  if (WhichCase == TD_CT) {
    double u00 = M11/4 + M12/4 + M13/2 + M14/4 + M22/12 + M23/4 + M24/6 + M33/4
                 + M34/4 + M44/12 + (2*V1)/3 + V2/3 + (2*V3)/3 + V4/3 + S;
    U[0][0][T1] = u00;
    U[0][1][TY1] = M12/3 + M14/3 + M22/6 + M23/3 + M24/3 + M34/3 + M44/6 + V2/2 + V4/2;
    U[0][1][T1]  = -(2*M11)/3 - (5*M12)/6 - (4*M13)/3 - (5*M14)/6 - M22/3 - (5*M23)/6
                   - (2*M24)/3 - (2*M33)/3 - (5*M34)/6 - M44/3 - (3*V1)/2 - V2 - (3*V3)/2
                   - V4 - 2*S;
    U[0][2][TY12] = M22/4 + M44/4;
    U[0][2][TY1]  = -M12/2 - M14 - M22/2 - M23 - M24 - M34/2 - M44/2 - V2 - V4;
    U[0][2][T1]   = (3*M11)/4 + M12 + M13 + M14 + M22/2 + M23 + M24
                    + (3*M33)/4 + M34 + M44/2 + V1 + V2 + V3 + V4 + S;
    U[0][3][TY12] = -M22/2 - M44/2;
    U[0][3][TY1]  = M14 + M22/2 + M23 + M24 + M44/2 + V2/2 + V4/2;
    U[0][3][T1]   = -M11/2 - M12/2 - M14/2 - M22/3 - M23/2 - (2*M24)/3
                    - M33/2 - M34/2 - M44/3 - V1/6 - V2/3 - V3/6 - V4/3;
    U[0][4][TY12] = M22/4 + M44/4;
    U[0][4][TY1]  = M12/6 - M14/3 - M22/6 - M23/3 - M24/3 + M34/6 - M44/6;
    U[0][4][T1]   = M11/6 + M12/12 - M13/6 + M14/12 + M22/12 + M23/12 + M24/6
                    + M33/6 + M34/12 + M44/12;

    U[1][0][T1] = u00;
    U[1][1][TY1] = -M11/3 - M12/2 - (2*M13)/3 - M14/2 - M22/6 - M23/2 - M24/3 - M33/3
                   - M34/2 - M44/6 - V1/2 - V2/2 - V3/2 - V4/2;
    U[1][1][T1]  = -M11/3 - M12/3 - (2*M13)/3 - M14/3 - M22/6 - M23/3 - M24/3 - M33/3
                   - M34/3 - M44/6 - 2*S - V1 - V2/2 - V3 - V4/2;
    U[1][2][TY12] = M11/4 + M12/2 + M22/4 + M33/4 + M34/2 + M44/4;
    U[1][2][TY1]  = M11/2 + M12/2 + M13 + M14 + M23 + M24 + M33/2 + M34/2
                    + V1 + V2 + V3 + V4;
    U[1][2][T1]   = M22/4 + M44/4 + S;
    U[1][3][TY12] = -M11/2 - M12 - M22/2 - M33/2 - M34 - M44/2;
    U[1][3][TY1]  = M12/2 - M14/2 + M22/2 - M23/2 - M24 + M34/2 + M44/2
                    - V1/2 - V2/2 - V3/2 - V4/2;
    U[1][3][T1]   = -M22/3 + M24/3 - M44/3 + V1/3 + V2/6 + V3/3 + V4/6;
    U[1][4][TY12] = M11/4 + M12/2 + M22/4 + M33/4 + M34/2 + M44/4;
    U[1][4][TY1]  = -M11/6 - M12/2 - M13/3 - M22/3 + M24/3 - M33/6 - M34/2 - M44/3;
    U[1][4][T1]   = M11/12 + M12/12 + M13/6 + M14/12 + M22/6 + M23/12 - M24/6
                    + M33/12 + M34/12 + M44/6;

    U[2][0][T1] = M11/4 + M12/4 + M13/2 + M14/4 + M22/12 + M23/4 + M24/6 + M33/4
                  + M34/4 + M44/12 + S + (2*V1)/3 + V2/3 + (2*V3)/3 + V4/3;
    U[2][1][TY1] = -M11/3 - M12/6 - (2*M13)/3 - M14/6 - M23/6 - M33/3 - M34/6 - V1/2 - V3/2;
    U[2][1][T1]  = -M11/3 - M12/3 - (2*M13)/3 - M14/3 - M22/6 - M23/3 - M24/3
                   - M33/3 - M34/3 - M44/6 - 2*S - V1 - V2/2 - V3 - V4/2;
    U[2][2][TY12] = M11/4 + M33/4;
    U[2][2][TY1]  = M11/2 + M12/2 + M13 + M33/2 + M34/2 + V1 + V3;
    U[2][2][T1]   = M22/4 + M44/4 + S;
    U[2][3][TY12] = -M11/2 - M33/2;
    U[2][3][TY1]  = -M12/2 + M14/2 + M23/2 - M34/2 - V1/2 - V3/2;
    U[2][3][T1]   = -M22/3 + M24/3 - M44/3 + V1/3 + V2/6 + V3/3 + V4/6;
    U[2][4][TY12] = M11/4 + M33/4;
    U[2][4][TY1]  = -M11/6 + M12/6 - M13/3 - M14/3 - M23/3 - M33/6 + M34/6;
    U[2][4][T1]   = M11/12 + M12/12 + M13/6 + M14/12 + M22/6 + M23/12 - M24/6
                    + M33/12 + M34/12 + M44/6;
  } else if (WhichCase == TD_CE) {
    double u00 = C2010/4 + M11/6 + M13/3 + M33/6 + S + V1/2 + V3/2;
    double u10 = C2001/3 + M12/2 + M14/2 + M23/2 + M34/2 - S + V2 + V4;
    double u20 = C0210/2 + M22/2 + M24 + M44/2 - V1/2 - V2 - V3/2 - V4;
    double u30 = C0201 - M11/6 - M12/2 - M13/3 - M14/2 - M22/2 - M23/2 - M24 - M33/6 - M34/2 - M44/2;
    double u40 = -C0201 - C0210/2 - C2001/3 - C2010/4;

    /* d = 0 */
    U[0][0][T1] = u00;
    U[0][1][TY1Y2] = M12/2 + M23/2 + V2;
    U[0][1][TY1]   = -C2001/3 - C2010/3 - M12/2 - M13/2 - M14/2 - M23/2 - M33/2 - M34/2 - V2 - V3 - V4;
    U[0][1][T1]    = u10;
    U[0][2][TY12Y22] = C0210/2 + M22/2;
    U[0][2][TY12Y2]  = -C0210 - M22 - M23 - M24;
    U[0][2][TY12]    = C0210/2 + M22/2 + M23 + M24 + M33/2 + M34 + M44/2;
    U[0][2][TY1Y2]   = C0210 + M22 + M24 - V2;
    U[0][2][TY1]     = -C0210 - M22 - M23 - 2*M24 - M34 - M44 + V2 + V3 + V4;
    U[0][2][T1]      = u20;
    U[0][3][TY13Y22] = -C0201 - C0210;
    U[0][3][TY13Y2]  = 2*C0201 + 2*C0210;
    U[0][3][TY13]    = -C0201 - C0210;
    U[0][3][TY12Y22] = C0201 - M22/2;
    U[0][3][TY12Y2]  = -4*C0201 - 2*C0210 + M22 + M23 + M24;
    U[0][3][TY12]    = 3*C0201 + 2*C0210 - M22/2 - M23 - M24 - M33/2 - M34 - M44/2;
    U[0][3][TY1Y2]   = 2*C0201 - M12/2 - M22 - M23/2 - M24;
    U[0][3][TY1]     = -3*C0201 - C0210 + M12/2 + M13/2 + M14/2 + M22 + (3*M23)/2 + 2*M24
                       + M33/2 + (3*M34)/2 + M44;
    U[0][3][T1]      = u30;
    U[0][4][TY13Y22] = C0201 + C0210;
    U[0][4][TY13Y2]  = -2*C0201 - 2*C0210;
    U[0][4][TY13]    = C0201 + C0210;
    U[0][4][TY12Y22] = -C0201 - C0210/2;
    U[0][4][TY12Y2]  = 4*C0201 + 3*C0210;
    U[0][4][TY12]    = -3*C0201 - (5*C0210)/2;
    U[0][4][TY1Y2]   = -2*C0201 - C0210;
    U[0][4][TY1]     = 3*C0201 + 2*C0210 + C2001/3 + C2010/3;
    U[0][4][T1]      = u40;

    /* d = 1 */
    U[1][0][T1] = u00;
    U[1][1][TY1Y2] = C2001/3 + M14/2 + M34/2 + V4;
    U[1][1][TY1]   = -C2001/3 - (2*C2010)/3 - M11/2 - M12/2 - M13/2 - M14/2 - M23/2 - M34/2 - V1 - V2 - V4;
    U[1][1][T1]    = u10;
    U[1][2][TY12Y22] = M44/2;
    U[1][2][TY12Y2]  = -C2001 - M14 - M24 - M44;
    U[1][2][TY12]    = C0210/2 + C2001 + C2010/2 + M11/2 + M12 + M14 + M22/2 + M24 + M44/2;
    U[1][2][TY1Y2]   = M24 + M44 - V4;
    U[1][2][TY1]     = -C0210 - C2001 - M12 - M14 - M22 - 2*M24 - M44 + V1 + V2 + V4;
    U[1][2][T1]      = u20;
    U[1][3][TY13Y2]  = C0201 + C2001;
    U[1][3][TY13]    = -C0201 - C2001;
    U[1][3][TY12Y22] = -M44/2;
    U[1][3][TY12Y2]  = -2*C0201 + M14 + M24 + M44;
    U[1][3][TY12]    = 3*C0201 + C2001 - M11/2 - M12 - M14 - M22/2 - M24 - M44/2;
    U[1][3][TY1Y2]   = C0201 - M14/2 - M24 - M34/2 - M44;
    U[1][3][TY1]     = -3*C0201 + M11/2 + (3*M12)/2 + M13/2 + (3*M14)/2 + M22 + M23/2 + 2*M24
                       + M34/2 + M44;
    U[1][3][T1]      = u30;
    U[1][4][TY13Y2]  = -C0201 - C2001;
    U[1][4][TY13]    = C0201 + C2001;
    U[1][4][TY12Y2]  = 2*C0201 + C2001;
    U[1][4][TY12]    = -3*C0201 - C0210/2 - 2*C2001 - C2010/2;
    U[1][4][TY1Y2]   = -C0201 - C2001/3;
    U[1][4][TY1]     = 3*C0201 + C0210 + (4*C2001)/3 + (2*C2010)/3;
    U[1][4][T1]      = u40;

    /* d = 2 */
    U[2][0][T1] = u00;
    U[2][1][TY1Y2] = -C2001/3 - C2010/3 - M13/2 - M14/2 - M33/2 - M34/2 - V3 - V4;
    U[2][1][TY1]   = -M12/2 - M23/2 - V2;
    U[2][1][T1]    = u10;
    U[2][2][TY12Y22] = M33/2 + M34 + M44/2;
    U[2][2][TY12Y2]  = M23 + M24;
    U[2][2][TY12]    = C0210/2 + M22/2;
    U[2][2][TY1Y2]   = -M23 - M24 - M34 - M44 + V3 + V4;
    U[2][2][TY1]     = -C0210 - M22 - M24 + V2;
    U[2][2][T1]      = u20;
    U[2][3][TY13Y2]  = -C0201 - C0210;
    U[2][3][TY12Y22] = -M33/2 - M34 - M44/2;
    U[2][3][TY12Y2]  = 2*C0201 + 2*C0210 - M23 - M24;
    U[2][3][TY12]    = C0201 - M22/2;
    U[2][3][TY1Y2]   = -C0201 - C0210 + M13/2 + M14/2 + M23 + M24 + M33/2 + (3*M34)/2 + M44;
    U[2][3][TY1]     = -2*C0201 + M12/2 + M22 + M23/2 + M24;
    U[2][3][T1]      = u30;
    U[2][4][TY13Y2]  = C0201 + C0210;
    U[2][4][TY12Y2]  = -2*C0201 - 2*C0210;
    U[2][4][TY12]    = -C0201 - C0210/2;
    U[2][4][TY1Y2]   = C0201 + C0210 + C2001/3 + C2010/3;
    U[2][4][TY1]     = 2*C0201 + C0210;
    U[2][4][T1]      = u40;

    /* d = 3 */
    U[3][0][T1] = u00;
    U[3][1][TY1Y2] = (-2*C2010)/3 - M11/2 - M12/2 - M13/2 - M23/2 - V1 - V2;
    U[3][1][TY1]   = -C2001/3 - M14/2 - M34/2 - V4;
    U[3][1][T1]    = u10;
    U[3][2][TY12Y22] = C0210/2 + C2010/2 + M11/2 + M12 + M22/2;
    U[3][2][TY12Y2]  = C2001 + M14 + M24;
    U[3][2][TY12]    = M44/2;
    U[3][2][TY1Y2]   = -C0210 - C2001 - M12 - M14 - M22 - M24 + V1 + V2;
    U[3][2][TY1]     = -M24 - M44 + V4;
    U[3][2][T1]      = u20;
    U[3][3][TY13Y22] = -C0201 - C2001;
    U[3][3][TY12Y22] = C0201 + C2001 - M11/2 - M12 - M22/2;
    U[3][3][TY12Y2]  = 2*C0201 - M14 - M24;
    U[3][3][TY12]    = -M44/2;
    U[3][3][TY1Y2]   = -2*C0201 + M11/2 + (3*M12)/2 + M13/2 + M14 + M22 + M23/2 + M24;
    U[3][3][TY1]     = -C0201 + M14/2 + M24 + M34/2 + M44;
    U[3][3][T1]      = u30;
    U[3][4][TY13Y22] = C0201 + C2001;
    U[3][4][TY12Y22] = -C0201 - C0210/2 - C2001 - C2010/2;
    U[3][4][TY12Y2]  = -2*C0201 - C2001;
    U[3][4][TY1Y2]   = 2*C0201 + C0210 + C2001 + (2*C2010)/3;
    U[3][4][TY1]     = C0201 + C2001/3;
    U[3][4][T1]      = u40;

    /* d = 4 */
    U[4][0][T1] = u00;
    U[4][1][TY1Y2] = -C2010/3 - M13/2 - M33/2 - V3;
    U[4][1][TY1]   = -C2001/3 - M14/2 - M34/2 - V4;
    U[4][1][T1]    = u10;
    U[4][2][TY12Y22] = M33/2;
    U[4][2][TY12Y2]  = M34;
    U[4][2][TY12]    = M44/2;
    U[4][2][TY1Y2]   = -M23 - M34 + V3;
    U[4][2][TY1]     = -M24 - M44 + V4;
    U[4][2][T1]      = u20;
    U[4][3][TY12Y22] = -M33/2;
    U[4][3][TY12Y2]  = -M34;
    U[4][3][TY12]    = -M44/2;
    U[4][3][TY1Y2]   = -C0210 + M13/2 + M23 + M33/2 + M34;
    U[4][3][TY1]     = -C0201 + M14/2 + M24 + M34/2 + M44;
    U[4][3][T1]      = u30;
    U[4][4][TY1Y2]   = C0210 + C2010/3;
    U[4][4][TY1]     = C0201 + C2001/3;
    U[4][4][T1]      = u40;

    /* d = 5 */
    U[5][0][T1] = u00;
    U[5][1][TY1Y2] = (-2*C2010)/3 - M11/2 - M13/2 - V1;
    U[5][1][TY1]   = -M12/2 - M23/2 - V2;
    U[5][1][T1]    = u10;
    U[5][2][TY12Y22] = C2010/2 + M11/2;
    U[5][2][TY12Y2]  = M12;
    U[5][2][TY12]    = C0210/2 + M22/2;
    U[5][2][TY1Y2]   = -C2001 - M12 - M14 + V1;
    U[5][2][TY1]     = -C0210 - M22 - M24 + V2;
    U[5][2][T1]      = u20;
    U[5][3][TY12Y22] = C2001 - M11/2;
    U[5][3][TY12Y2]  = -M12;
    U[5][3][TY12]    = C0201 - M22/2;
    U[5][3][TY1Y2]   = M11/2 + M12 + M13/2 + M14;
    U[5][3][TY1]     = -2*C0201 + M12/2 + M22 + M23/2 + M24;
    U[5][3][T1]      = u30;
    U[5][4][TY12Y22] = -C2001 - C2010/2;
    U[5][4][TY12]    = -C0201 - C0210/2;
    U[5][4][TY1Y2]   = C2001 + (2*C2010)/3;
    U[5][4][TY1]     = 2*C0201 + C0210;
    U[5][4][T1]      = u40;
  } else {
    /* common vertex */
    U[0][0][T1] = S;
    U[0][1][T1] = V1; U[0][1][TY1] = V2; U[0][1][TY2] = V3; U[0][1][TY2Y3] = V4;
    U[0][2][T1] = M11/2.0; U[0][2][TY1] = M12; U[0][2][TY12] = M22/2.0;
    U[0][2][TY2] = M13; U[0][2][TY1Y2] = M23; U[0][2][TY22] = M33/2.0;
    U[0][2][TY2Y3] = M14; U[0][2][TY1Y2Y3] = M24; U[0][2][TY22Y3] = M34; U[0][2][TY22Y32] = M44/2.0;
    U[0][3][TY2] = C2010; U[0][3][TY12Y2] = C0210; U[0][3][TY2Y3] = C2001; U[0][3][TY12Y2Y3] = C0201;

    U[1][0][T1] = S;
    U[1][1][T1] = V3; U[1][1][TY1] = V4; U[1][1][TY2] = V1; U[1][1][TY2Y3] = V2;
    U[1][2][T1] = M33/2.0; U[1][2][TY1] = M34; U[1][2][TY12] = M44/2.0;
    U[1][2][TY2] = M13; U[1][2][TY1Y2] = M14; U[1][2][TY22] = M11/2.0;
    U[1][2][TY2Y3] = M23; U[1][2][TY1Y2Y3] = M24; U[1][2][TY22Y3] = M12; U[1][2][TY22Y32] = M22/2.0;
    U[1][3][TY22] = C2010; U[1][3][TY1Y22] = C2001; U[1][3][TY22Y32] = C0210; U[1][3][TY1Y22Y32] = C0201;
  }
}


// distance functions X_d(y) of the once-integrable kernels (TaylorDuffy.cc:506-566)
__device__ inline double ce_x(const TDGeo &W, double u1, double u2, double xi2)
{
  return sqrt(u1 * u1 * W.A2 + u2 * u2 * W.BP2 + xi2 * xi2 * W.L2
              + 2.0 * (u1 * (u2 * W.AdBP + xi2 * W.AdL) + u2 * xi2 * W.BPdL));
}
__device__ inline double cv_x(const TDGeo &W, double xi1, double xi2, double eta1, double eta2)
{
  return sqrt(xi1 * xi1 * W.A2 + xi2 * xi2 * W.B2 + eta1 * eta1 * W.AP2 + eta2 * eta2 * W.BP2
              + 2.0 * xi1 * (xi2 * W.AdB - eta1 * W.AdAP - eta2 * W.AdBP)
              - 2.0 * xi2 * (eta1 * W.BdAP + eta2 * W.BdBP)
              + 2.0 * eta1 * eta2 * W.APdBP);
}
__device__ void td_x(const TDGeo &W, const double *y, double *X)
{
  if (W.WhichCase == TD_CT) {
    double t = y[0], u1, u2;
    u1 = 1.0; u2 = t;     X[0] = sqrt(W.A2 * u1 * u1 + 2.0 * W.AdB * u1 * u2 + W.B2 * u2 * u2);
    u1 = t; u2 = (t - 1); X[1] = sqrt(W.A2 * u1 * u1 + 2.0 * W.AdB * u1 * u2 + W.B2 * u2 * u2);
    u1 = t; u2 = 1.0;     X[2] = sqrt(W.A2 * u1 * u1 + 2.0 * W.AdB * u1 * u2 + W.B2 * u2 * u2);
  } else if (W.WhichCase == TD_CE) {
    double y1 = y[0], y2 = y[1];
    X[0] = ce_x(W, -y1, -y1 * y2, 1.0 - y1 + y1 * y2);
    X[1] = ce_x(W, y1, y1 * y2, 1.0 - y1);
    X[2] = ce_x(W, -y1 * y2, y1 * (1.0 - y2), 1.0 - y1);
    X[3] = ce_x(W, y1 * y2, -y1 * (1.0 - y2), 1.0 - y1 * y2);
    X[4] = ce_x(W, -y1 * y2, -y1, 1.0);
    X[5] = ce_x(W, y1 * y2, y1, 1.0 - y1);
  } else {
    double y1 = y[0], y2 = y[1], y3 = y[2];
    X[0] = cv_x(W, 1.0, y1, y2, y2 * y3);
    X[1] = cv_x(W, y2, y2 * y3, 1.0, y1);
  }
}

// script P of one polynomial at y (TaylorDuffy.cc:568-634): P[d][n][0..1]
__device__ void td_script_p(const TDGeo &W, const double (*U)[TD_NWP][TD_NMON], int nMin, int nMax,
                            const double *y, double (*P)[TD_NWP][2])
{
  if (W.WhichCase == TD_CT) {
    for (int d = 0; d < 3; d++)
      for (int n = nMin; n <= nMax; n++) { P[d][n][0] = U[d][n][T1]; P[d][n][1] = U[d][n][TY1]; }
  } else if (W.WhichCase == TD_CE) {
    double y1 = y[0], y12 = y1 * y1;
    for (int d = 0; d < 6; d++)
      for (int n = nMin; n <= nMax; n++) {
        P[d][n][0] = U[d][n][T1] + U[d][n][TY1] * y1 + U[d][n][TY12] * y12;
        P[d][n][1] = U[d][n][TY2] + U[d][n][TY1Y2] * y1 + U[d][n][TY12Y2] * y12;
      }
  } else {
    double y1 = y[0], y12 = y1 * y1, y2 = y[1], y22 = y2 * y2;
    for (int d = 0; d < 2; d++)
      for (int n = nMin; n <= nMax; n++) {
        P[d][n][0] = U[d][n][T1] + U[d][n][TY1] * y1 + U[d][n][TY2] * y2 + U[d][n][TY12] * y12
                     + U[d][n][TY1Y2] * y1 * y2 + U[d][n][TY22] * y22 + U[d][n][TY12Y2] * y12 * y2
                     + U[d][n][TY1Y22] * y1 * y22;
        P[d][n][1] = U[d][n][TY2Y3] * y2 + U[d][n][TY1Y2Y3] * y1 * y2 + U[d][n][TY22Y3] * y22
                     + U[d][n][TY12Y2Y3] * y12 * y2;
      }
  }
}

// relative exponential ExpRelV3P0 (TaylorDuffy.cc:840-893)
__device__ cplx sw_exprel(int n, cplx Z)
{
  int m;
  if (Z.i == 0.0) {
    double X = Z.r, Term, Sum;
    if (fabs(X) < 0.1) {
      Sum = 1.0;
      for (Term = 1.0, m = 1; m < 100; m++) {
        Term *= X / ((double)(m + n));
        Sum += Term;
        if (fabs(Term) < 1.0e-8 * fabs(Sum)) break;
      }
      return mk(Sum, 0);
    }
    Sum = exp(X);
    for (Term = 1.0, m = 0; m < n; m++) { Sum -= Term; Term *= X / ((double)(m + 1)); }
    return mk(Sum / Term, 0);
  }
  cplx Term, Sum;
  if (sqrt(norm2(Z)) < 0.1) {
    Sum = Term = mk(1, 0);
    for (m = 1; m < 100; m++) {
      Term = Term * (Z / (double)(m + n));
      Sum += Term;
      if (norm2(Term) < 1.0e-16 * norm2(Sum)) break;
    }
    return Sum;
  }
  Sum = cexp_(Z);
  Term = mk(1, 0);
  for (m = 0; m < n; m++) { Sum -= Term; Term = Term * (Z / (double)(m + 1)); }
  return Sum / Term;
}

// script K of the Helmholtz kernels (TaylorDuffy.cc:896-935)
__device__ void td_script_k(int WhichK, cplx KParam, double X, int nMin, int nMax, cplx *K)
{
  cplx IK = IMUL(KParam), IKX = IK * X, e = cexp_(IKX);
  if (WhichK == K_HELM) {
    for (int n = nMin; n <= nMax; n++) K[n] = e * sw_exprel(n, -IKX) / (n * X);
  } else {
    cplx ER[10];
    if ((nMin - 2) < 0 || (nMax - 1) >= 10) {
      for (int n = 0; n < nMin; n++) K[n] = mk(0, 0);
      nMin = 2;
    }
    for (int n = nMin - 2; n <= nMax - 1; n++) ER[n] = sw_exprel(n, -IKX);
    for (int n = nMin; n <= nMax; n++)
      K[n] = e * (IK * ER[n - 1] / ((double)n - 1.0) - ER[n - 2] / (((double)n - 2.0) * X)) / (X * X);
  }
}

// TaylorDuffySum (TaylorDuffy.cc:152-258) for the once-integrable
// (Helmholtz) kernels: f[2 npk], f[2 npk + 1]
__device__ void td_sum(const TDGeo &W, const double (*U)[TD_NREG][TD_NWP][TD_NMON], const double *y, double *f)
{
  int NumRegions, nOffset;
  double yIn, Jac;
  if (W.WhichCase == TD_CT) { NumRegions = 3; nOffset = 1; yIn = y[0]; Jac = 1.0; }
  else if (W.WhichCase == TD_CE) { NumRegions = 6; nOffset = 2; yIn = y[1]; Jac = y[0]; }
  else { NumRegions = 2; nOffset = 3; yIn = y[2]; Jac = y[1]; }
  if (Jac == 0.0) { for (int q = 0; q < 2 * W.NumPKs; q++) f[q] = 0.0; return; }
  double X[TD_NREG];
  td_x(W, y, X);
  for (int npk = 0; npk < W.NumPKs; npk++) {
    const int nMin = W.nMin[npk], nMax = W.nMax[npk];
    double P[TD_NREG][TD_NWP][2];
    td_script_p(W, U[npk], nMin, nMax, y, P);
    cplx K[TD_NREG][8];
    for (int d = 0; d < NumRegions; d++) td_script_k(W.KIndex[npk], W.k, X[d], nMin + nOffset, nMax + nOffset, K[d]);
    cplx s = mk(0, 0);
    for (int n = nMin; n <= nMax; n++)
      for (int d = 0; d < NumRegions; d++) s += (P[d][n][0] + yIn * P[d][n][1]) * K[d][n + nOffset];
    s = s * (Jac / (4.0 * M_PI));
    f[2 * npk] = s.r; f[2 * npk + 1] = s.i;
  }
}

// ComputeGeometricParameters (TaylorDuffy.cc:1265-1528) for P = UNITY,
// PMCHWG1, PMCHWC (the first NumPKs of them): the geometric scalars into W
// (every lane), and, on lane npk < NumPKs, the Upsilon table of polynomial
// npk into U[npk] and its lowest and highest nonzero power into nm[npk],
// nm[TD_NPK + npk]
__device__ void td_geometry(int WhichCase, const double *V1, const double *V2, const double *V3, const double *V2P,
                            const double *V3P, const double *Q, const double *QP, TDGeo &W,
                            double (*U)[TD_NREG][TD_NWP][TD_NMON], int *nm)
{
  if (WhichCase == TD_CT) { V2P = V2; V3P = V3; }
  else if (WhichCase == TD_CE) V2P = V2;
  double A[3], B[3], AP[3], BP[3], L[3];
  sub3(V2, V1, A); sub3(V3, V2, B); sub3(V2P, V1, AP); sub3(V3P, V2P, BP);
  W.A2 = dot3(A, A); W.AdB = dot3(A, B); W.B2 = dot3(B, B);
  W.AP2 = W.BP2 = W.L2 = W.AdAP = W.AdBP = W.AdL = W.BdAP = W.BdBP = W.APdBP = W.BPdL = 0.0;
  if (WhichCase == TD_CE) {
    sub3(BP, B, L);
    W.BP2 = dot3(BP, BP); W.AdBP = dot3(A, BP); W.L2 = dot3(L, L);
    W.AdL = dot3(A, L); W.BPdL = dot3(BP, L);
  } else if (WhichCase == TD_CV) {
    W.AP2 = dot3(AP, AP); W.BP2 = dot3(BP, BP); W.AdAP = dot3(A, AP); W.AdBP = dot3(A, BP);
    W.BdAP = dot3(B, AP); W.BdBP = dot3(B, BP); W.APdBP = dot3(AP, BP);
  }
  const int npk = threadIdx.x;
  if (npk >= W.NumPKs) return;
  const int NumRegions = (WhichCase == TD_CV) ? 2 : (WhichCase == TD_CE ? 6 : 3);
  double D[3], DP[3], DeltaD[3], Scr[3], Scr2[3];
  sub3(V1, Q, D); sub3(V1, QP, DP); sub3(QP, Q, DeltaD);
  double C[4] = {0, 0, 0, 0}, M[16], V[4] = {0, 0, 0, 0}, S = 0.0;
  for (int q = 0; q < 16; q++) M[q] = 0.0;
  if (npk == 0) S = 1;
  else if (npk == 1) {          // PMCHWG1
    M[0 + 2 * 4] = dot3(A, AP); M[0 + 3 * 4] = dot3(A, BP);
    M[1 + 2 * 4] = dot3(B, AP); M[1 + 3 * 4] = dot3(B, BP);
    V[0] = dot3(A, DP); V[1] = dot3(B, DP); V[2] = dot3(D, AP); V[3] = dot3(D, BP);
    S = dot3(D, DP);
  } else {                      // PMCHWC
    cross3(A, AP, Scr); M[0 + 2 * 4] = -dot3(Scr, DeltaD);
    cross3(A, BP, Scr); M[0 + 3 * 4] = -dot3(Scr, DeltaD);
    cross3(B, AP, Scr); M[1 + 2 * 4] = -dot3(Scr, DeltaD);
    cross3(B, BP, Scr); M[1 + 3 * 4] = -dot3(Scr, DeltaD);
    cross3(DP, D, Scr2);
    V[0] = -dot3(A, Scr2); V[1] = -dot3(B, Scr2); V[2] = dot3(AP, Scr2); V[3] = dot3(BP, Scr2);
  }
  td_upsilon(WhichCase, C, M, V, S, U[npk]);
  int nMin = TD_NWP - 1, nMax = -1;
  for (int n = 0; n < TD_NWP; n++)
    for (int nr = 0; nr < NumRegions; nr++)
      for (int q = 0; q < TD_NMON; q++)
        if (fabs(U[npk][nr][n][q]) > 1.0e-8) { if (n < nMin) nMin = n; if (n > nMax) nMax = n; }
  nm[npk] = nMin; nm[TD_NPK + npk] = nMax;
}

// ---------------------------------------------------------------------
// p-adaptive Clenshaw-Curtis cubature on [0,1]^dim, dim <= 3, with the
// value cache and traversal of pcubature.c:60-430, run by one wave: the
// points of a new block are evaluated across the lanes, and the rule sums
// over the cache (the full rule and its reduction in each dimension) are
// taken in one lane-strided pass and summed over the wave.  Cache entry e
// holds the points of rule orders m[] whose dimension mi is at its newest
// level (mi = dim: the full initial 3^dim grid), dimension 0 slowest.
// ---------------------------------------------------------------------
struct PCEntry { int m0, m1, m2, mi, off; };

__device__ inline int pc_count(int m, bool isnew) { return isnew ? (m ? (1 << m) : 2) : (1 << (m + 1)) + 1; }

// coordinate of list index l in one dimension (c = r = 1/2)
__device__ inline double pc_coord(int m, bool isnew, int l)
{
  if (isnew) {
    int j = l >> 1;
    double x = c_swx[(m ? (1 << (m - 1)) : 0) + j];
    return (l & 1) ? 0.5 - 0.5 * x : 0.5 + 0.5 * x;
  }
  if (l == 0) return 0.5;
  int j = (l - 1) >> 1;
  return ((l - 1) & 1) ? 0.5 - 0.5 * c_swx[j] : 0.5 + 0.5 * c_swx[j];
}

// weight of list index l of a cached dimension (cm, isnew) in the rule of
// order m reduced by one when red; < 0 when the value does not take part
__device__ inline double pc_weight(int cm, bool isnew, int m, bool red, int l)
{
  if (m == 0 && red) return l == 0 ? 2.0 : -1.0;
  const int mid = m - (red ? 1 : 0);
  const double *w = c_sww + mid + (1 << mid) - 1;
  if (isnew) {
    const int off = cm ? 1 + (1 << (cm - 1)) : 1;
    return w[off + (l >> 1)];
  }
  if (l == 0) return w[0];
  const int j = (l - 1) >> 1;
  const int nx = cm <= mid ? (1 << cm) : (1 << mid);
  return j < nx ? w[1 + j] : -1.0;
}

// sum over the 64 lanes; every lane gets the same bits (each butterfly
// step adds the same two operands on both partners)
__device__ inline double wave_sum(double v)
{
  for (int s = 32; s >= 1; s >>= 1) v += __shfl_xor(v, s, 64);
  return v;
}

// appends the block of points of rule orders m[] whose dimension mi is at
// its newest level: entry E[ne] (LDS, lane 0), values at vals + nval
template <typename F>
__device__ void pc_add(F &f, int dim, const int *m, int mi, PCEntry *E, int &ne, double *vals, int &nval)
{
  if (threadIdx.x == 0) {
    PCEntry &e = E[ne];
    e.m0 = m[0]; e.m1 = dim > 1 ? m[1] : 0; e.m2 = dim > 2 ? m[2] : 0; e.mi = mi; e.off = nval;
  }
  ne++;
  int cnt[3] = {1, 1, 1};
  for (int d = 0; d < dim; d++) cnt[d] = pc_count(m[d], d == mi);
  const int total = cnt[0] * cnt[1] * cnt[2];
  for (int L = threadIdx.x; L < total; L += 64) {
    int l[3] = {0, 0, 0}, rem = L;
    for (int d = dim - 1; d >= 0; d--) { l[d] = rem % cnt[d]; rem /= cnt[d]; }
    double y[3];
    for (int d = 0; d < dim; d++) y[d] = pc_coord(m[d], d == mi, l[d]);
    f(y, vals + (long)(nval + L) * SW_FDIM);
  }
  nval += total;
  __syncthreads();
}

// the rule of order m over the cache, reduced in dimension md (md = dim:
// none), for md = 0..dim at once: S[md][q], identical on every lane
__device__ void pc_evals(int dim, const PCEntry *E, int ne, const double *vals, const int *m, double V,
                         double (*S)[SW_FDIM])
{
#pragma unroll
  for (int md = 0; md < 4; md++)
#pragma unroll
    for (int q = 0; q < SW_FDIM; q++) S[md][q] = 0.0;
  for (int c = 0; c < ne; c++) {
    const PCEntry e = E[c];
    const int cm[3] = {e.m0, e.m1, e.m2};
    bool inc[4];
#pragma unroll
    for (int md = 0; md < 4; md++)
      inc[md] = md <= dim && (e.mi >= dim || cm[e.mi] + (e.mi == md) <= m[e.mi]);
    if (!(inc[0] || inc[1] || inc[2] || inc[3])) continue;
    int cnt[3] = {1, 1, 1};
    for (int d = 0; d < dim; d++) cnt[d] = pc_count(cm[d], d == e.mi);
    const int total = cnt[0] * cnt[1] * cnt[2];
    for (int L = threadIdx.x; L < total; L += 64) {
      int l[3] = {0, 0, 0}, rem = L;
      for (int d = dim - 1; d >= 0; d--) { l[d] = rem % cnt[d]; rem /= cnt[d]; }
      double wf[3] = {1, 1, 1}, wr[3] = {1, 1, 1};
#pragma unroll
      for (int d = 0; d < 3; d++)
        if (d < dim) {
          wf[d] = pc_weight(cm[d], d == e.mi, m[d], false, l[d]);
          wr[d] = pc_weight(cm[d], d == e.mi, m[d], true, l[d]);
        }
      const double *v = vals + (long)(e.off + L) * SW_FDIM;
      double vv[SW_FDIM];
#pragma unroll
      for (int q = 0; q < SW_FDIM; q++) vv[q] = v[q];
#pragma unroll
      for (int md = 0; md < 4; md++) {
        if (!inc[md]) continue;
        double wt = V;
        bool use = true;
#pragma unroll
        for (int d = 0; d < 3; d++)
          if (d < dim) {
            const double w = d == md ? wr[d] : wf[d];
            if (w < 0) use = false;
            else wt *= w;
          }
        if (!use) continue;
#pragma unroll
        for (int q = 0; q < SW_FDIM; q++) S[md][q] += vv[q] * wt;
      }
    }
  }
#pragma unroll
  for (int md = 0; md < 4; md++)
    if (md <= dim)
#pragma unroll
      for (int q = 0; q < SW_FDIM; q++) S[md][q] = wave_sum(S[md][q]);
}

// the integral into val[SW_FDIM] (every lane); vals: the job's value
// cache in scratch memory, E: its entries in LDS
template <typename F>
__device__ void pcubature_wave(F &f, int dim, double *vals, PCEntry *E, double *val)
{
  int m[3] = {0, 0, 0};
  double V = 1.0;
  for (int d = 0; d < dim; d++) V *= 0.5;
  int ne = 0, nval = 0;
  long numEval = 0;
  pc_add(f, dim, m, dim, E, ne, vals, nval);
  for (;;) {
    double S[4][SW_FDIM];
    pc_evals(dim, E, ne, vals, m, V, S);
    double err[SW_FDIM];
#pragma unroll
    for (int q = 0; q < SW_FDIM; q++) {
      val[q] = dim == 1 ? S[1][q] : (dim == 2 ? S[2][q] : S[3][q]);
      err[q] = 0.0;
    }
    int mi = 0;
    double maxerr = 0.0;
#pragma unroll
    for (int i = 0; i < 3; i++) {
      if (i >= dim) continue;
      double emax = 0.0;
#pragma unroll
      for (int q = 0; q < SW_FDIM; q++) {
        double e = fabs(val[q] - S[i][q]);
        if (e > emax) emax = e;
        if (e > err[q]) err[q] = e;
      }
      if (emax > maxerr) { maxerr = emax; mi = i; }
    }
    bool conv = true;
#pragma unroll
    for (int q = 0; q < SW_FDIM; q++)
      if (err[q] > 0.0 && err[q] > fabs(val[q]) * TD_RELTOL) conv = false;
    if (conv || numEval > TD_MAXEVAL) break;
    m[mi] += 1;
    if (m[mi] >= PC_M) break;     // not reached within TD_MAXEVAL points
    int nb = 1;
    for (int d = 0; d < dim; d++) nb *= pc_count(m[d], d == mi);
    if (ne >= PC_MAXENT || nval + nb > PC_MAXVAL) { m[mi] -= 1; break; }
    pc_add(f, dim, m, mi, E, ne, vals, nval);
    numEval += nb;
  }
}

// ---------------------------------------------------------------------
// the jobs
// ---------------------------------------------------------------------
// 9 (Qa, Qb) combinations from panel-pair moments in panel-local
// coordinates (the same recombination as the far and near kernels)
__device__ void sw_combos(const cplx S1, const cplx *Sx, const cplx *Sy, const cplx Sxy, const cplx s, const cplx *u,
                          const cplx *v, const cplx *w, const double *Qa, const double *Qb, cplx k, int needH1,
                          cplx *H0, cplx *H1)
{
  cplx ik = IMUL(k);
  cplx c4 = 4.0 / (ik * ik);
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) {
      const double *qa = Qa + 3 * a, *qb = Qb + 3 * b;
      cplx h = Sxy;
      for (int c = 0; c < 3; c++) h = h - qb[c] * Sx[c] - qa[c] * Sy[c];
      h += (mk(qa[0] * qb[0] + qa[1] * qb[1] + qa[2] * qb[2], 0) + c4) * S1;
      H0[3 * a + b] = h;
      if (needH1) {
        double qq[3];
        cross3(qa, qb, qq);
        cplx t = s;
        for (int c = 0; c < 3; c++) t = t - qb[c] * u[c] - qa[c] * v[c] + qq[c] * w[c];
        H1[3 * a + b] = t;
      }
    }
}

// no common vertex: 78 x 78-point rule without desingularization
// (GetPPIs_Cubature with HighOrder, PanelPanelInteractions.cc:437-441);
// the 6084 point pairs strided over the lanes, the moments summed over the
// wave (every lane gets them)
__device__ void sw_hocub(const NearPair &P, cplx k, int needH1, cplx *H0, cplx *H1)
{
  double aA[3], aB[3], bA[3], bB[3], D[3], Qal[9], Qbl[9];
  for (int c = 0; c < 3; c++) {
    aA[c] = P.Va[3 + c] - P.Va[c]; aB[c] = P.Va[6 + c] - P.Va[c];
    bA[c] = P.Vb[3 + c] - P.Vb[c]; bB[c] = P.Vb[6 + c] - P.Vb[c];
    D[c] = P.Va[c] - P.Vb[c];
    for (int q = 0; q < 3; q++) { Qal[3 * q + c] = P.Qa[3 * q + c] - P.Va[c]; Qbl[3 * q + c] = P.Qb[3 * q + c] - P.Vb[c]; }
  }
  // S1, Sx[3], Sy[3], Sxy, s, u[3], v[3], w[3]
  cplx acc[18];
  for (int q = 0; q < 18; q++) acc[q] = mk(0, 0);
  cplx ik = IMUL(k);
  for (int p = threadIdx.x; p < 78 * 78; p += 64) {
    const int i = p / 78, j = p - 78 * (p / 78);
    const double wi = c_swtcr20[3 * i + 2], wj = c_swtcr20[3 * j + 2];
    double x[3], y[3];
    for (int c = 0; c < 3; c++) {
      x[c] = c_swtcr20[3 * i] * aA[c] + c_swtcr20[3 * i + 1] * aB[c];
      y[c] = c_swtcr20[3 * j] * bA[c] + c_swtcr20[3 * j + 1] * bB[c];
    }
    double R[3] = {D[0] + x[0] - y[0], D[1] + x[1] - y[1], D[2] + x[2] - y[2]};
    double r = sqrt(R[0] * R[0] + R[1] * R[1] + R[2] * R[2]), ir = 1.0 / r;
    double e = exp(-k.i * r), sn, cs;
    sincos(k.r * r, &sn, &cs);
    cplx Phi = mk(e * cs, e * sn) * (SCUFFX_INV4PI * ir) * (wi * wj);
    acc[0] += Phi;
    for (int c = 0; c < 3; c++) { acc[1 + c] += x[c] * Phi; acc[4 + c] += y[c] * Phi; }
    acc[7] += (x[0] * y[0] + x[1] * y[1] + x[2] * y[2]) * Phi;
    if (needH1) {
      cplx Psi = Phi * (ik - mk(ir, 0)) * ir;
      double xy[3], Rx[3], yR[3];
      cross3(x, y, xy); cross3(R, x, Rx); cross3(y, R, yR);
      acc[8] += (xy[0] * R[0] + xy[1] * R[1] + xy[2] * R[2]) * Psi;
      for (int c = 0; c < 3; c++) { acc[9 + c] += Rx[c] * Psi; acc[12 + c] += yR[c] * Psi; acc[15 + c] += R[c] * Psi; }
    }
  }
  for (int q = 0; q < (needH1 ? 18 : 8); q++) { acc[q].r = wave_sum(acc[q].r); acc[q].i = wave_sum(acc[q].i); }
  sw_combos(acc[0], acc + 1, acc + 4, acc[7], acc[8], acc + 9, acc + 12, acc + 15, Qal, Qbl, k, needH1, H0, H1);
}

static_assert(SW_SLAB == SCUFFX_SW_SLAB, "scratch slab size");

struct TDFun {
  const TDGeo *W;
  const double (*U)[TD_NREG][TD_NWP][TD_NMON];
  __device__ void operator()(const double *y, double *out) const { td_sum(*W, U, y, out); }
};

// one wave (block of 64) per job; kind 0: no-common-vertex near pair (all
// 9 combinations), 1: common-edge/vertex near pair, one (Qa, Qb).  A
// Taylor-Duffy job keeps its Upsilon tables and cache entries in LDS and
// its cached integrand values in a slab of scratch memory.
__global__ void __launch_bounds__(64)
k_sw(const SWJob *jobs, int job0, int njobs, const NearPair *nps, const double *V, const int *PVI, int nreg,
     const RegionK *regs, int needH1, double *HN, double *HS, double *scratch)
{
  __shared__ double sU[TD_NPK][TD_NREG][TD_NWP][TD_NMON];
  __shared__ PCEntry sE[PC_MAXENT];
  __shared__ int sNm[2 * TD_NPK];
  const int t = blockIdx.x;
  if (t >= njobs) return;
  const SWJob J = jobs[job0 + t];
  const cplx k = regs[J.region].k;
  const int stride = needH1 ? 4 : 2;
  const int lane = threadIdx.x;
  if (J.kind == 0) {
    cplx H0[9], H1[9];
    sw_hocub(nps[J.pair], k, needH1, H0, H1);
    double *o = HN + ((long long)J.pair * nreg + J.region) * 9 * stride;
    if (lane == 0)
      for (int q = 0; q < 9; q++) {
        o[q * stride] = H0[q].r; o[q * stride + 1] = H0[q].i;
        if (needH1) { o[q * stride + 2] = H1[q].r; o[q * stride + 3] = H1[q].i; }
      }
    return;
  }
  double *vals = scratch + (long long)t * SW_SLAB;
  TDGeo W;
  W.k = k;
  const int a = J.q / 3, b = J.q % 3;
  const NearPair &P = nps[J.pair];
  W.WhichCase = P.ncv; W.NumPKs = 3;
  W.KIndex[0] = K_HELM; W.KIndex[1] = K_HELM; W.KIndex[2] = K_GRADHELM;
  td_geometry(P.ncv, P.Va, P.Va + 3, P.Va + 6, P.Vb + 3, P.Vb + 6, P.Qa + 3 * a, P.Qb + 3 * b, W, sU, sNm);
  __syncthreads();
  for (int npk = 0; npk < TD_NPK; npk++) { W.nMin[npk] = sNm[npk]; W.nMax[npk] = sNm[TD_NPK + npk]; }
  TDFun f{&W, sU};
  double R[SW_FDIM];
  pcubature_wave(f, 4 - P.ncv, vals, sE, R);
  if (lane == 0) {
    cplx H0 = mk(R[2], R[3]) - 4.0 * mk(R[0], R[1]) / (k * k), H1 = mk(R[4], R[5]);
    double *o = HN + ((long long)J.pair * nreg + J.region) * 9 * stride + J.q * stride;
    o[0] = H0.r; o[1] = H0.i;
    if (needH1) { o[2] = H1.r; o[3] = H1.i; }
  }
}

static bool g_sw_ready[64] = {false};

hipError_t launch_sw(hipStream_t stream, const SWJob *jobs, int njobs, const NearPair *nps, const double *V,
                     const int *PVI, int nreg, const RegionK *regs, int needH1, double *HN, double *HS,
                     double *scratch, int chunk)
{
  if (njobs <= 0) return hipSuccess;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return hipErrorInvalidDevice;
  if (!g_sw_ready[dev & 63]) {
    // Clenshaw-Curtis nodes and weights in the nested layout of
    // pcubature.c: x[0] = 1; level m >= 1 at offset 2^(m-1); weights of
    // rule order m (2^(m+1) intervals) at offset m + 2^m - 1:
    // [center, level-0 node, level-1 nodes, ...]
    static double hx[PC_NX], hw[PC_NW];
    const long double pi = 3.141592653589793238462643383279502884L;
    hx[0] = 1.0;
    for (int m = 1; m < PC_M; m++)
      for (int i = 0; i < (1 << (m - 1)); i++)
        hx[(1 << (m - 1)) + i] = (double)cosl((2 * i + 1) * pi / (long double)(1L << (m + 1)));
    auto wt = [&](long n, long j) {
      long double s = 0;
      for (long q = 1; q <= n / 2; q++) {
        long double b = (q == n / 2) ? 1.0L : 2.0L;
        s += b / (4.0L * q * q - 1.0L) * cosl(2.0L * pi * (long double)((j * q) % n) / (long double)n);
      }
      long double cj = (j == 0 || j == n) ? 1.0L : 2.0L;
      return (double)(cj / (long double)n * (1.0L - s));
    };
    for (int m = 0; m < PC_M; m++) {
      long n = 1L << (m + 1);
      double *w = hw + m + (1 << m) - 1;
      w[0] = wt(n, n / 2);
      w[1] = wt(n, 0);
      int q = 2;
      for (int l = 1; l <= m; l++)
        for (int i = 0; i < (1 << (l - 1)); i++) w[q++] = wt(n, (2 * i + 1) * (n >> (l + 1)));
    }
    hipError_t e = hipMemcpyToSymbol(HIP_SYMBOL(c_swx), hx, sizeof(hx));
    if (e == hipSuccess) e = hipMemcpyToSymbol(HIP_SYMBOL(c_sww), hw, sizeof(hw));
    if (e == hipSuccess) e = hipMemcpyToSymbol(HIP_SYMBOL(c_swtcr20), scuffx_tcr20, sizeof(scuffx_tcr20));
    if (e != hipSuccess) return e;
    g_sw_ready[dev & 63] = true;
  }
  for (int j0 = 0; j0 < njobs; j0 += chunk) {
    int n = std::min(chunk, njobs - j0);
    hipLaunchKernelGGL(k_sw, dim3(n), dim3(64), 0, stream, jobs, j0, n, nps, V, PVI, nreg, regs, needH1,
                       HN, HS, scratch);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace scuffx
