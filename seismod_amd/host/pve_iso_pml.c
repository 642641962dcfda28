/*
 * pve_iso_pml -- drop-in for seismod's pve_iso_pml: poroviscoelastic wave
 * propagation in 2D isotropic media on an MI355X (pve_driver.c).
 */
#include "pve_driver.h"

int main(int argc, char *argv[])
{
    return sm_pve_program(argc, argv, 0);
}
