/* ref_dkt_entry.c -- entry point of the DKT test harness: starts the Fortran
 * runtime as a Fortran main program would, then runs ref_dkt_run(nsteps).
 * TEST INFRASTRUCTURE ONLY.  usage: ref_dkt nsteps */
#include <stdlib.h>

void _FortranAProgramStart(int argc, const char *argv[], const char *envp[], const void *env_defaults);
void _FortranAProgramEndStatement(void);
void ref_dkt_run(int nsteps);

int main(int argc, const char *argv[], const char *envp[]) {
    _FortranAProgramStart(argc, argv, envp, 0);
    ref_dkt_run(argc > 1 ? atoi(argv[1]) : 1);
    _FortranAProgramEndStatement();
    return 0;
}
