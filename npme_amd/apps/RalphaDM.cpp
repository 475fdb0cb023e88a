// RalphaDM coordChargeFile keywordFile -- the command-line tool of
// app/RalphaDM.cpp:44-216 (funcType Ralpha, EwaldSplit DerivMatch; alpha from the keyword file).
#include "calc_tool.hpp"

int main(int argc, char **argv) { return npme_tools::run_calc(argc, argv, "Ralpha"); }
