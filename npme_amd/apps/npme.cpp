// npme coordChargeFile keywordFile -- the command-line tool of
// app/npme.cpp:41-462 (funcType and EwaldSplit from the keyword file).
#include "calc_tool.hpp"

int main(int argc, char **argv) { return npme_tools::run_calc(argc, argv, nullptr); }
