/* A libm shim for conditioning tests: loaded with LD_PRELOAD, it returns the
 * neighbouring double (one ulp up or down) of pow or sincos on about half of
 * their arguments, chosen by a hash of the arguments, for the functions named
 * in the PERTURB environment variable.  Everything else passes through. */
#define _GNU_SOURCE
#include <dlfcn.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static int active(const char *name)
{
   const char *e = getenv("PERTURB");
   return e && strstr(e, name);
}

static uint64_t mix(double a, double b)
{
   uint64_t x, y;
   memcpy(&x, &a, 8);
   memcpy(&y, &b, 8);
   x ^= y * 0x9E3779B97F4A7C15ull;
   x ^= x >> 29;
   x *= 0xBF58476D1CE4E5B9ull;
   x ^= x >> 32;
   return x;
}

static double nudge(double r, uint64_t k)
{
   if (!(k & 1)) return r;
   return nextafter(r, (k & 2) ? INFINITY : -INFINITY);
}

double pow(double a, double b)
{
   static double (*f)(double, double);
   if (!f) f = (double (*)(double, double))dlsym(RTLD_NEXT, "pow");
   double r = f(a, b);
   return active("pow") ? nudge(r, mix(a, b)) : r;
}

void sincos(double a, double *s, double *c)
{
   static void (*f)(double, double *, double *);
   if (!f) f = (void (*)(double, double *, double *))dlsym(RTLD_NEXT, "sincos");
   f(a, s, c);
   if (active("sincos")) {
      *s = nudge(*s, mix(a, 3.0));
      *c = nudge(*c, mix(a, 4.0));
   }
}
