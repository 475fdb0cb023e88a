// laplaceDM coordChargeFile keywordFile -- the command-line tool of
// app/laplaceDM.cpp:44-191 (funcType Laplace, EwaldSplit DerivMatch).
#include "calc_tool.hpp"

int main(int argc, char **argv) { return npme_tools::run_calc(argc, argv, "Laplace"); }
