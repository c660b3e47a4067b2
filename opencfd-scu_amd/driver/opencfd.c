/* opencfd.c - single-process driver with the control flow of the reference's
 * main (opencfd.c:23-53): read the input, allocate, initialise, run the time
 * loop, release.  MPI is not used: the library runs on one GPU. */
#include "ocfd.h"

int main(void) {
    read_parameters();
    opencfd_mem_init_all();
    init();
    NS_solver_real();
    opencfd_mem_finalize_all();
    return 0;
}
