// main.cpp -- chfem_exec command line (chfem/src/main.c:33-298): same flags,
// same order of hmg* calls through runAnalysis.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/chfem_amd.h"
#include "../../include/femhmg.h"

static void print_help()
{
    printf("\nComputational Homogenization with the image-based FEM on MI355X\n\n");
    printf("./chfem_exec [.nf] [.raw] ... (options)\n\nor\n\n./chfem_exec (options) ... -i [.nf] [.raw] ... (options)\n\n");
    printf("\t-b: Save results in a binary file. Must be followed by a string with a filename.\n");
    printf("\t-c: Stopping criteria for the PCG method: 0 - L2 (default), 1 - Inf, 2 - L2+Error.\n");
    printf("\t-d: Target direction: 0 - X, 1 - Y, 2 - Z, 3 - YZ, 4 - XZ, 5 - XY, 6 - ALL (default).\n");
    printf("\t-e: Export fields from simulation (by nodes).\n");
    printf("\t-f: Input scalar density field. Must be followed by a [.bin] file.\n");
    printf("\t-h: Print this help info and exit.\n");
    printf("\t-i: Input files. Must be followed by: [.nf] [.raw].\n");
    printf("\t-j: Jacobi preconditioning: 0 - no, 1 - yes (default).\n");
    printf("\t-m: Write metrics report. Must be followed by a string with a filename.\n");
    printf("\t-p: Parallel matrix-free strategy: 0 - NBN (default), 1 - EBE.\n");
    printf("\t-pm: Pore mapping strategy: 0 - image, 1 - DOF number (default).\n");
    printf("\t-prec: Arithmetic: 0 - FP64 (default), 1 - FP32 inner solve with FP64 refinement.\n");
    printf("\t-r: Number of recursive searches for initial guesses.\n");
    printf("\t-s: Solver: 0 - CG (default), 1 - MINRES, 2 - CG3, 3 - MINRES3, 4 - CG2, 5 - MINRES2.\n");
    printf("\t-u: Reduce strategy for velocity fields (FLUID): 0 - on the fly, 1 - only diag, 2 - full.\n");
    printf("\t-xi: Import initial guesses for PCG solver from binary files. Must be followed by a string with a filename.\n");
    printf("\t-xo: Export result vector (x) from the PCG solver.\n\n");
}

static int find_flag(const char *flag, char *arr[], int sz)
{
    for (int i = 1; i < sz; i++)
        if (!strcmp(arr[i], flag)) return i;
    return 0;
}

static char *entry(int i, char *arr[], int sz)
{
    if (i >= sz) {
        printf("Invalid input after %s.\nCall -h for help with input directives.\n", arr[sz - 1]);
        return NULL;
    }
    return arr[i];
}

static int read_input(char *arr[], int sz, chfemgpuInput_t *u, unsigned *prec)
{
    if (sz < 2) {
        printf("Invalid input.\nCall -h for help with input directives.\n");
        return 0;
    }
    if (find_flag("-h", arr, sz)) {
        print_help();
        return 2;
    }
    if (sz == 2) {
        printf("Invalid input ignored: %s.\nCall -h for help with input directives.\n", arr[1]);
        return 0;
    }
    int id = find_flag("-i", arr, sz);
    u->neutral_file = entry(id + 1, arr, sz);
    u->raw_file = entry(id + 2, arr, sz);
    int i = 1 + (!id) * 2;
    while (i < sz) {
        if (i != id) {
            unsigned *prop = NULL;
            int valid = 0;
            if (!strcmp(arr[i], "-r")) prop = &u->num_of_recursions_initguess;
            else if (!strcmp(arr[i], "-j")) prop = &u->preconditioner_flag;
            else if (!strcmp(arr[i], "-u")) prop = &u->xreduce_flag;
            else if (!strcmp(arr[i], "-s")) prop = &u->solver_flag;
            else if (!strcmp(arr[i], "-p")) prop = &u->parallel_flag;
            else if (!strcmp(arr[i], "-pm")) prop = &u->poremap_flag;
            else if (!strcmp(arr[i], "-d")) prop = &u->hmg_direction_flag;
            else if (!strcmp(arr[i], "-c")) prop = &u->stopcrit_flag;
            else if (!strcmp(arr[i], "-prec")) prop = prec;
            else if (!strcmp(arr[i], "-e")) { u->exportFields_flag = 1; valid = 1; }
            else if (!strcmp(arr[i], "-b")) {
                char *s = entry(++i, arr, sz); if (!s) return 0;
                u->save2binary_flag = 1; u->binary_file = s; valid = 1;
            } else if (!strcmp(arr[i], "-f")) {
                char *s = entry(++i, arr, sz); if (!s) return 0;
                u->sdf_bin_file = s; valid = 1;
            } else if (!strcmp(arr[i], "-m")) {
                char *s = entry(++i, arr, sz); if (!s) return 0;
                u->writeReport_flag = 1; u->report_file = s; valid = 1;
            } else if (!strcmp(arr[i], "-xi")) {
                char *s = entry(++i, arr, sz); if (!s) return 0;
                u->importX_flag = 1; u->x0_file = s; valid = 1;
            } else if (!strcmp(arr[i], "-xo")) { u->exportX_flag = 1; valid = 1; }
            if (prop) {
                char *s = entry(++i, arr, sz); if (!s) return 0;
                *prop = (unsigned)atoi(s);
            } else if (!valid) printf("Invalid input ignored: %s.\n", arr[i]);
        } else i += 2;
        i++;
    }
    int ok = 1;
    if (!u->neutral_file || !u->raw_file) { printf("Invalid input, .nf and .raw files are required to characterize model.\n"); ok = 0; }
    if (u->xreduce_flag > 2) { printf("Invalid reduce strategy flag.\n"); ok = 0; }
    if (u->solver_flag > 5) { printf("Invalid solver flag.\n"); ok = 0; }
    if (u->parallel_flag > 1) { printf("Invalid parallel strategy flag.\n"); ok = 0; }
    if (!ok) printf("Call -h for help with input directives.\n");
    return ok;
}

int main(int argc, char *argv[])
{
    chfemgpuInput_t u;
    initDefaultInput(&u);
    unsigned prec = 0;
    int rc = read_input(argv, argc, &u, &prec);
    if (!rc) { printf("Failed to parse arguments.\nProcess aborted.\n"); return 0; }
    if (rc == 2) return 0;
    printf("#######################################################\n");
    if (!hmgInit(u.neutral_file, u.raw_file, u.sdf_bin_file, NULL)) {
        printf("Failed to properly read input files.\nProcess aborted.\n#######################################################\n");
        return -1;
    }
    hmgSetPrecisionFlag((cudapcgFlag_t)prec);
    if (runAnalysis(&u) != 0) {
        printf("Failed to run analysis.\nProcess aborted.\n#######################################################\n");
        return -1;
    }
    return 0;
}
