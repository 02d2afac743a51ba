/*
 * options.c -- a small options database with PETSc's command-line syntax
 * (the reference reads every setting through PetscOptionsGet*; the defaults
 * and names are those of se2wave.c, listed at each call site of se2w_main).
 *
 * Syntax: "-name value" or "-name" (bool true); booleans also accept
 * true/false/yes/no/1/0; arrays are comma-separated ("-src_xy 1,2,3,4").
 * A token is an option name when it starts with '-' followed by a letter, so
 * negative numbers are values.  A later occurrence overrides an earlier one.
 */
#include <ctype.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "host_internal.h"

struct se2w_options {
  int n;
  char **name;  /* without the leading '-' */
  char **value; /* NULL for a bare flag */
};

static int is_option_token(const char *t) { return t && t[0] == '-' && isalpha((unsigned char)t[1]); }

static char *dupstr(const char *s)
{
  const size_t n = strlen(s) + 1;
  char *d = (char *)malloc(n);
  if (d) memcpy(d, s, n);
  return d;
}

int se2w_options_create(int argc, char **argv, SE2WOptions *out)
{
  struct se2w_options *o;
  int i;
  if (!out) return host_fail(SE2W_HOST_ERR_ARG, "null output");
  o = (struct se2w_options *)calloc(1, sizeof(*o));
  if (!o) return host_fail(SE2W_HOST_ERR_MEM, "out of memory");
  o->name = (char **)calloc(argc > 0 ? argc : 1, sizeof(char *));
  o->value = (char **)calloc(argc > 0 ? argc : 1, sizeof(char *));
  if (!o->name || !o->value) { se2w_options_destroy(&o); return host_fail(SE2W_HOST_ERR_MEM, "out of memory"); }
  for (i = 1; i < argc; i++) {
    if (!is_option_token(argv[i])) continue;
    o->name[o->n] = dupstr(argv[i] + 1);
    if (i + 1 < argc && !is_option_token(argv[i + 1])) {
      o->value[o->n] = dupstr(argv[i + 1]);
      i++;
    }
    o->n++;
  }
  *out = o;
  return 0;
}

int se2w_options_destroy(SE2WOptions *op)
{
  int i;
  if (!op || !*op) return 0;
  for (i = 0; i < (*op)->n; i++) { free((*op)->name[i]); free((*op)->value[i]); }
  free((*op)->name); free((*op)->value);
  free(*op);
  *op = NULL;
  return 0;
}

/* last occurrence of -name, or -1 */
static int lookup(SE2WOptions o, const char *name)
{
  int i, hit = -1;
  if (!o || !name) return -1;
  if (name[0] == '-') name++;
  for (i = 0; i < o->n; i++)
    if (strcmp(o->name[i], name) == 0) hit = i;
  return hit;
}

int se2w_options_get_int(SE2WOptions o, const char *name, int *value, int *found)
{
  const int i = lookup(o, name);
  char *end;
  long v;
  if (found) *found = 0;
  if (i < 0) return 0;
  if (!o->value[i]) return host_fail(SE2W_HOST_ERR_ARG, "option -%s needs an integer value", o->name[i]);
  v = strtol(o->value[i], &end, 10);
  if (*end) return host_fail(SE2W_HOST_ERR_ARG, "option -%s: '%s' is not an integer", o->name[i], o->value[i]);
  *value = (int)v;
  if (found) *found = 1;
  return 0;
}

int se2w_options_get_real(SE2WOptions o, const char *name, double *value, int *found)
{
  const int i = lookup(o, name);
  char *end;
  double v;
  if (found) *found = 0;
  if (i < 0) return 0;
  if (!o->value[i]) return host_fail(SE2W_HOST_ERR_ARG, "option -%s needs a real value", o->name[i]);
  v = strtod(o->value[i], &end);
  if (*end) return host_fail(SE2W_HOST_ERR_ARG, "option -%s: '%s' is not a real number", o->name[i], o->value[i]);
  *value = v;
  if (found) *found = 1;
  return 0;
}

int se2w_options_get_bool(SE2WOptions o, const char *name, int *value, int *found)
{
  const int i = lookup(o, name);
  const char *v;
  if (found) *found = 0;
  if (i < 0) return 0;
  v = o->value[i];
  if (!v || !strcmp(v, "1") || !strcasecmp(v, "true") || !strcasecmp(v, "yes")) *value = 1;
  else if (!strcmp(v, "0") || !strcasecmp(v, "false") || !strcasecmp(v, "no")) *value = 0;
  else return host_fail(SE2W_HOST_ERR_ARG, "option -%s: '%s' is not a boolean", o->name[i], v);
  if (found) *found = 1;
  return 0;
}

int se2w_options_get_string(SE2WOptions o, const char *name, char *value, size_t len, int *found)
{
  const int i = lookup(o, name);
  if (found) *found = 0;
  if (i < 0) return 0;
  if (!o->value[i]) return host_fail(SE2W_HOST_ERR_ARG, "option -%s needs a value", o->name[i]);
  snprintf(value, len, "%s", o->value[i]);
  if (found) *found = 1;
  return 0;
}

int se2w_options_get_real_array(SE2WOptions o, const char *name, double *values, int *n, int *found)
{
  const int i = lookup(o, name);
  const char *p;
  int cap, cnt = 0;
  if (found) *found = 0;
  if (!n) return host_fail(SE2W_HOST_ERR_ARG, "null count");
  cap = *n;
  if (i < 0) return 0;
  if (!o->value[i]) return host_fail(SE2W_HOST_ERR_ARG, "option -%s needs a list of reals", o->name[i]);
  p = o->value[i];
  while (*p) {
    char *end;
    const double v = strtod(p, &end);
    if (end == p) return host_fail(SE2W_HOST_ERR_ARG, "option -%s: bad list '%s'", o->name[i], o->value[i]);
    if (cnt >= cap) return host_fail(SE2W_HOST_ERR_ARG, "option -%s: more than %d values", o->name[i], cap);
    values[cnt++] = v;
    p = end;
    if (*p == ',') p++;
    else if (*p) return host_fail(SE2W_HOST_ERR_ARG, "option -%s: bad list '%s'", o->name[i], o->value[i]);
  }
  *n = cnt;
  if (found) *found = 1;
  return 0;
}
