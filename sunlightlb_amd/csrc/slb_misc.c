/*
 * slb_misc.c -- runtime helpers of the sunlightlb.h API: fatal errors,
 * verbosity, parameter and status reports, the monitor file, string helpers
 * and USR1/USR2 signal handling (reference src/lbmisc.c:39-378,
 * src/lbsignal.c:40-88).  Report texts keep the reference's formats so a
 * driver's output reads the same.
 */
#include <signal.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <unistd.h>
#include "slb_internal.h"

int slb_verbosity = 0;
#ifdef SIGNAL
int slb_signal_flag = 0;
#endif

static time_t tstart;

void slb_clock_reset(void) { tstart = 0; }

/* Non-recoverable error: message to stderr and exit (src/lbmisc.c:39-42). */
void error(char *s1) {
  fprintf(stderr, "Opps!  %s\n", s1);
  exit(EXIT_FAILURE);
}

void startclock() {
  if (tstart == 0) printf("Clock started\n");
  else printf("Clock restarted\n");
  tstart = time(NULL);
}

void set_verbosity(int what) {
  slb_verbosity |= what;
  report_verbosity();
}

void unset_verbosity(int what) {
  if (slb_verbosity & what) slb_verbosity -= what;
  report_verbosity();
}

void report_verbosity() {
  printf("Verbosity =");
  if (slb_verbosity == 0) { printf(" NONE\n"); return; }
  if (slb_verbosity & TRACE_ROUTINE_EXITS) printf(" TRACE_ROUTINE_EXITS");
  if (slb_verbosity & OBJECT_DETAILS) printf(" OBJECT_DETAILS");
  if (slb_verbosity & FORCE_DETAILS) printf(" FORCE_DETAILS");
  printf("\n");
}

void about() {
  printf("\nSunlightLB-compatible 3D lattice Boltzmann library for AMD Instinct MI355X\n");
  printf("D3Q15 stress-relaxation LB, link bounce-back and moment propagation on\n");
  printf("hand-written gfx950 HIP kernels behind the SunlightLB C API.\n");
  printf("GPU: %s\n\n", slbd_device_count() > 0 ? slbd_arch() : "none visible");
}

void wrlinks(LBSys lb) {
  for (int l = 0; l < 15; l++)
    printf(" w%02i = %2i  c%02i = (%2i,%2i,%2i)\n", l, lb->w[l], l, lb->c[3 * l],
           lb->c[3 * l + 1], lb->c[3 * l + 2]);
  printf("\n");
}

/* Parameter report (src/lbmisc.c:193-242).  The expected total force
 * includes the live uniform body force dpdx/dpdy/dpdz. */
void wrpars(LBSys lb) {
  int nobjs = 0, nfluid = 0;
  double fxsum = 0.0, fysum = 0.0, fzsum = 0.0;
  for (LBObj p = lb->obj_head; p != NULL; p = p->next_obj) nobjs++;
  for (long n = 0; n < lb->nnodes; n++)
    if (lb->obj[n] == NULL) {
      fxsum += lb->bx[n]; fysum += lb->by[n]; fzsum += lb->bz[n];
      nfluid++;
    }
  fxsum += nfluid * (double)lb->dpdx;
  fysum += nfluid * (double)lb->dpdy;
  fzsum += nfluid * (double)lb->dpdz;
  LBReal nu = ((lb->tau) - 0.5) / 3.0, eta = (lb->rho0) * nu;
  long msize = (long)(24 * sizeof(LBReal) + sizeof(LBObj) + sizeof(int)) * (lb->nnodes);
  printf("---------------------------------------------------\n");
  printf(" Operating parameters\n");
  printf("---------------------------------------------------\n");
  printf(" LB size = %i * %i * %i = %i\n", lb->esi, lb->esj, lb->esk, lb->nnodes);
  if (sizeof(LBReal) == sizeof(double)) printf(" Using double precision\n");
  else printf(" Using single precision\n");
  printf(" Memory requirement approx %0.2f MByte\n", (double)(msize) / (1024 * 1024));
  printf(" rho0 = %0.4f\n", lb->rho0);
  printf(" tau, nu, eta = %0.4f %0.4f %0.4f\n", lb->tau, nu, eta);
  printf(" Number of objects = %i\n", nobjs);
  printf(" Number of fluid + solid nodes = %i + %i = %i\n", nfluid, (lb->nnodes) - nfluid,
         (lb->nnodes));
  printf(" Expected total Fx, y, z = %0.2f %0.2f %0.2f\n", fxsum, fysum, fzsum);
  if (lb->stokes_flow) printf(" Stokes equations simulated\n");
  else printf(" Navier-Stokes equations simulated\n");
  if (lb->vel_corrected) printf(" Corrected velocity field used\n");
  else printf(" Uncorrected velocity field used\n");
  if (lb->skip_solid) printf(" Relaxation only on fluid nodes\n");
  else printf(" Relaxation on fluid and solid nodes\n");
  if (lb->objects_moving) printf(" Objects allowed to move\n");
  else printf(" All objects are stationary\n");
  /* device details go to stderr: stdout keeps the reference's lines */
  fprintf(stderr, " GPU %s, %d-bit populations, %.2f MByte on device\n", slbd_arch(),
          PRIV(lb)->prec, (double)slb_device_bytes(lb) / (1024 * 1024));
#ifdef SIGNAL
  printf(" Signalling implemented, process id = %i\n", (int)getpid());
#else
  printf(" Signalling not implemented\n");
#endif
  if (lb->monitoring != 0) printf(" Monitoring data will be written to %s.mon\n", lb->name);
  else printf(" No monitoring\n");
  printf("---------------------------------------------------\n");
  if (slb_verbosity & TRACE_ROUTINE_EXITS) printf("LB %i : wrpars done\n", lb->step);
}

void wrstatus(LBSys lb, FILE *fp) {
  fprintf(fp, "Currently at ");
  switch (lb->status) {
    case POST_BBL: fprintf(fp, "post-bounceback"); break;
    case POST_PROPAGATE: fprintf(fp, "post-propagation"); break;
    case POST_COLLIDE: fprintf(fp, "post-collision"); break;
    default: fprintf(fp, "(help, I'm stymied!)"); break;
  }
  fprintf(fp, " stage in the LB update cycle\n");
  if (slb_verbosity & TRACE_ROUTINE_EXITS) printf("LB %i : wrstatus done\n", lb->step);
}

/* Elapsed time as "1 hour 2 mins 3 secs" (long) or "h:mm:ss" (short). */
static void wrhms(FILE *fp, const char *fmt, long secs) {
  if (secs < 0) return;
  long hours = secs / 3600, mins;
  secs -= 3600 * hours;
  mins = secs / 60;
  secs -= 60 * mins;
  if (fmt[0] != 'l') { fprintf(fp, "%li:%02li:%02li", hours, mins, secs); return; }
  if (hours == 0 && mins == 0 && secs == 0) { fprintf(fp, "no time at all"); return; }
  if (hours == 1) fprintf(fp, "1 hour ");
  else if (hours > 1) fprintf(fp, "%li hours ", hours);
  if (mins == 1) fprintf(fp, "1 min ");
  else if (mins > 1) fprintf(fp, "%li mins ", mins);
  if (secs == 0) fprintf(fp, "exactly");
  else if (secs == 1) fprintf(fp, "1 sec");
  else fprintf(fp, "%li secs", secs);
}

/* Run statistics; ksite-updates/sec counts every node (src/lbmisc.c:264-278). */
void fwrstats(FILE *fp, int lbstep, int lbsize) {
  double rate = 0.0;
  time_t telapse = time(NULL) - tstart;
  if (telapse > 0) rate = (double)lbstep / (double)(telapse);
  fprintf(fp, "LB step %6i: ", lbstep);
  wrhms(fp, "long", (long)telapse);
  fprintf(fp, " elapsed\n");
  fprintf(fp, " %6.2f updates/sec\n", rate);
  fprintf(fp, " %6.1f ksite-updates/sec\n", lbsize * rate / 1000.0);
}

void wrstats(LBSys lb) { fwrstats(stdout, lb->step, lb->nnodes); }

/* Monitor file <name>.mon (src/lbmisc.c:286-354). */
void wrmon(LBSys lb, char *what) {
  if (lb->monitoring == 0) return;
  char w0 = what[0];
  double rate = 0.0;
  time_t telapse = 0;
  char *monfile = mycat(lb->name, ".mon");
  FILE *fp = fopen(monfile, "w");
  if (fp == NULL) {
    fprintf(stderr, "wrmon: I couldn't open %s.\n", monfile);
    free(monfile);
    return;
  }
  switch (w0) {
    case 'i': case 's': startclock(); break;
    case 'r': case 'f': case 'e':
      telapse = time(NULL) - tstart;
      if (telapse > 0) rate = (double)(lb->step) / (double)(telapse);
      break;
  }
#ifdef SIGNAL
  int pid = (int)getpid();
  fprintf(fp, "%i %s ", pid, lb->name);
  switch (w0) {
    case 'i': case 's': fprintf(fp, "started 0:00:00 ? ?\n"); break;
    case 'r': case 'f': case 'e':
      fprintf(fp, w0 == 'r' ? "running " : "finished ");
      wrhms(fp, "short", (long)telapse);
      if (telapse > 0) fprintf(fp, " %0.2f %0.2f\n", rate, (lb->nnodes) * rate / 1000.0);
      else fprintf(fp, " ? ?\n");
      break;
  }
#endif
  switch (w0) {
    case 'r': case 'f': case 'e':
#ifndef SIGNAL
      fprintf(fp, "%s %s\n", lb->name, w0 == 'r' ? "running" : "finished");
#endif
      fwrstats(fp, lb->step, lb->nnodes);
      break;
  }
#ifdef SIGNAL
  if (w0 == 'i' || w0 == 's' || w0 == 'r')
    fprintf(fp, "To update this file type : kill -USR1 %i\n", pid);
#endif
  if (fclose(fp) != 0) {
    fprintf(stderr, "wrmon: I couldn't close %s.\n", monfile);
    exit(EXIT_FAILURE);
  }
  free(monfile);
  printf("Monitor file %s.mon %s at %i\n", lb->name, what, lb->step);
  if (slb_verbosity & TRACE_ROUTINE_EXITS) printf("LB %i : wrmon done\n", lb->step);
}

char *mycat(char *s, char *t) {
  char *v = (char *)slb_xmalloc(strlen(s) + strlen(t) + 1);
  strcpy(v, s);
  strcat(v, t);
  return v;
}

char *mycopy(char *s) { return mycat(s, ""); }

/* True when s is a prefix of t (src/lbmisc.c:375-378). */
int myscmp(char *s, char *t) { return strncmp(t, s, strlen(s)) == 0; }

#ifdef SIGNAL
/* Act on a pending signal at a safe point (end of a collide or MP step):
 * USR1 refreshes the monitor file, USR2 checkpoints <name>.dms and
 * <name>.obj (src/lbsignal.c:40-62). */
void sighandle(LBSys lb) {
  if (slb_signal_flag == 0) return;
  int flag = slb_signal_flag;
  slb_signal_flag = 0;
  switch (flag) {
    case 1:
      printf("Acting on -USR1 signal\n");
      if (lb->name != NULL) wrmon(lb, "running");
      break;
    case 2:
      printf("Acting on -USR2 signal\n");
      if (lb->name != NULL) {
        char *dmsfile = mycat(lb->name, ".dms");
        char *objfile = mycat(lb->name, ".obj");
        lbsave(lb, dmsfile);
        if (lb->obj_head != NULL) lbsaveobj(lb, objfile);
        free(dmsfile);
        free(objfile);
      }
      break;
  }
  if (slb_verbosity & TRACE_ROUTINE_EXITS) printf("LB %i : sighandle done\n", lb->step);
}

/* Signal handler: records which signal arrived (src/lbsignal.c:72-79). */
void sighook(int sig) {
  int flag = 0;
  switch (sig) {
    case SIGUSR1: flag = 1; signal(SIGUSR1, sighook); break;
    case SIGUSR2: flag = 2; signal(SIGUSR2, sighook); break;
  }
  if (slb_signal_flag == 0) slb_signal_flag = flag;
}

void siginit() {
  slb_signal_flag = 0;
  signal(SIGUSR1, sighook);
  signal(SIGUSR2, sighook);
}
#endif
