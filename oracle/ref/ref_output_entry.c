/* ref_output_entry.c -- entry point of the output-writer test harness:
 * starts the Fortran runtime as a Fortran main program would, then runs
 * ref_output_run(itc).  TEST INFRASTRUCTURE ONLY.  usage: ref_output itc */
#include <stdlib.h>

void _FortranAProgramStart(int argc, const char *argv[], const char *envp[], const void *env_defaults);
void _FortranAProgramEndStatement(void);
void ref_output_run(int itc);

int main(int argc, const char *argv[], const char *envp[]) {
    _FortranAProgramStart(argc, argv, envp, 0);
    ref_output_run(argc > 1 ? atoi(argv[1]) : 0);
    _FortranAProgramEndStatement();
    return 0;
}
