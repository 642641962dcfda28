/*
 * pve_vti_pml -- drop-in for seismod's pve_vti_pml: poroviscoelastic wave
 * propagation in 2D VTI media on an MI355X (pve_driver.c).
 */
#include "pve_driver.h"

int main(int argc, char *argv[])
{
    return sm_pve_program(argc, argv, 1);
}
