// calcir executable: the drop-in replacement of calcIR.exe.
#include "calcir.h"

int main(int argc, char** argv) { return cir_main(argc, argv); }
