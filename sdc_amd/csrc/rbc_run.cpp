// rbc_run.cpp -- the simulation executable (sdc_amd/rbc_run): the reference's main program
// (gpu_flow/code/main.cu) as a main() around rbc_run_main of libsdc_rbc.so.
#include "rbc_abi.h"

int main(int argc, char** argv) { return rbc_run_main(argc, (const char* const*)argv); }
