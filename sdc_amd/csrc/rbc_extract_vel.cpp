// rbc_extract_vel.cpp -- the state-file extraction executable (sdc_amd/rbc_extract_vel): the
// reference's main_extract_vel.cu as a main() around rbc_extract_vel_main of libsdc_rbc.so.
#include "rbc_abi.h"

int main(int argc, char** argv) { return rbc_extract_vel_main(argc, (const char* const*)argv); }
