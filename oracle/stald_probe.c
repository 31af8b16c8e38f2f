/* TEST INFRASTRUCTURE ONLY.  Exposes CVODES's stability-limit detector
 * (cvSLdet, ThirdParty/sundials/src/cvodes/cvodes.c:8256, a static function)
 * and the detector inputs of real integrations to the tests: this file
 * includes the reference's cvodes.c where it lies (oracle/Makefile builds it,
 * with the harness, into oracle/_ref/libamx_stald_ref.so in place of
 * cvodes.o), so that the detector of the GPU integrator can be compared with
 * CVODES's own on the same inputs.  Nothing of the product uses it. */
#include "cvodes.c"
#include "amx_oracle.h"

/* ss[i][k] (row-major [5][3]): i = 0 newest of the five steps, k = 0, 1, 2
 * for orders q-1, q, q+1 (cv_ssdat[i + 1][k + 1] of CVODES's 1-based array). */
int amx_oracle_sldet(const double* ss, int q)
{
  struct CVodeMemRec mem;
  memset(&mem, 0, sizeof(mem));
  mem.cv_q = q;
  for (int i = 0; i < 5; ++i)
    for (int k = 0; k < 3; ++k) mem.cv_ssdat[i + 1][k + 1] = ss[i * 3 + k];
  return cvSLdet(&mem);
}

static double* g_buf;
static int g_cap, g_n;

/* after a step: the detector's input of that step if cvBDFStab called the
 * detector (order >= 3, nscon >= q + 5 steps at the current step size) */
static void record_step(void* mem)
{
  CVodeMem cv = (CVodeMem)mem;
  if (cv->cv_q < 3 || cv->cv_nscon < cv->cv_q + 5) return;
  if (g_n < g_cap) {
    double* r = g_buf + (size_t)g_n * 16;
    r[0] = cv->cv_q;
    for (int i = 0; i < 5; ++i)
      for (int k = 0; k < 3; ++k) r[1 + i * 3 + k] = cv->cv_ssdat[i + 1][k + 1];
  }
  g_n++;
}

/* One forward integration (amx_oracle_trace_forward); writes up to cap
 * records of 16 doubles (q, ss[5][3]) and returns how many steps called the
 * detector, or the CVODES status when the integration failed. */
int amx_oracle_stald_samples(const amx_oracle_model* m, const amx_oracle_opts* o,
                             const double* theta, const double* kfix, int nt, const double* ts,
                             double* out, int cap)
{
  double dummy[9];
  g_buf = out;
  g_cap = cap;
  g_n = 0;
  amx_oracle_step_hook = record_step;
  const int r = amx_oracle_trace_forward(m, o, theta, kfix, nt, ts, dummy, 0);
  amx_oracle_step_hook = NULL;
  return r < 0 ? r : g_n;
}
