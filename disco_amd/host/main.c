/*
 * disco -- driver of one Disco run on the MI355X timestep (main.c:26-97 of the
 * reference, serial): reads ./in.par, sets the mesh up, then loops
 *     dt = getmindt(); disco_check_dt(&dt); disco_possibly_report(); timestep(dt);
 * until T_End or an `abort` file, appending disco_report() lines to report.dat
 * on possiblyOutput's schedule (and one for the final state), and writes
 * the final state to
 * output.dstate (DSTATE01 layout, disco_amd/state_io.py) and the Analysis
 * module's accumulated sums to output.diag (int32 nq, int32 Nr*Nz, double
 * t_avg, double Qrz[Nr*Nz*nq], as the reference keeps theTools).
 *
 * Usage: disco [max_steps [checkpoint]]
 *   max_steps   stop after this many steps in all (counting those before a
 *               restart); -1 runs to T_End
 *   checkpoint  an output.dstate of an earlier run to restart from
 */
#include <sys/time.h>

#include "disco_host.h"

static double wall(void)
{
   struct timeval tv;
   gettimeofday(&tv, NULL);
   return tv.tv_sec + 1e-6 * tv.tv_usec;
}

int disco_write_dstate(struct domain *d, const char *path);

int main(int argc, char **argv)
{
   long max_steps = argc > 1 ? atol(argv[1]) : -1;
   struct domain D;
   memset(&D, 0, sizeof D);
   if (disco_read_par(&D, "in.par")) return 1;
   if (disco_setup(&D)) return 1;
   if (argc > 2 && disco_restart_dstate(&D, argv[2])) return 1;
   if (argc <= 2) {
      FILE *rf = fopen("report.dat", "w");  /* main.c:64-67 */
      if (rf) fclose(rf);
   }
   remove("abort");
   double t0 = wall();
   while (!D.final_step && (max_steps < 0 || D.count_steps < max_steps)) {
      double dt = getmindt(&D);
      disco_check_dt(&D, &dt);
      disco_possibly_report(&D, 0);
      timestep(&D, dt);
   }
   double secs = wall() - t0;
   disco_possibly_report(&D, 1);
   long ncells = 0;
   for (int k = D.NgZa; k < D.Nz - D.NgZb; ++k)
      for (int j = D.NgRa; j < D.Nr - D.NgRb; ++j) ncells += D.Np[j + D.Nr * k];
   printf("{\"cells\": %ld, \"steps\": %d, \"t\": %.10e, \"seconds\": %.6f, "
          "\"zone_updates_per_s\": %.6e}\n",
          ncells, D.count_steps, D.t, secs,
          secs > 0 ? (double)ncells * D.count_steps / secs : 0.0);
   if (disco_write_dstate(&D, "output.dstate")) return 1;
   if (disco_sync_to_host(&D)) return 1;
   FILE *fq = fopen("output.diag", "wb");
   if (!fq) return 1;
   int qhdr[2] = {D.num_tools, D.Nr * D.Nz};
   fwrite(qhdr, sizeof(int), 2, fq);
   fwrite(&D.theTools.t_avg, sizeof(double), 1, fq);
   fwrite(D.theTools.Qrz, sizeof(double), (size_t)qhdr[0] * qhdr[1], fq);
   fclose(fq);
   disco_free(&D);
   return 0;
}
