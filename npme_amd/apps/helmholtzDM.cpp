// helmholtzDM coordChargeFile keywordFile -- the command-line tool of
// app/helmholtzDM.cpp:44-191 (funcType Helmholtz, EwaldSplit DerivMatch; complex charges).
#include "calc_tool.hpp"

int main(int argc, char **argv) { return npme_tools::run_calc(argc, argv, "Helmholtz"); }
