// cylinder_fem_amr command-line program (the reference's executable, src/cylinder_fem_amr.cu:119).
extern "C" int cfa_main(int argc, char** argv);

int main(int argc, char** argv) { return cfa_main(argc, argv); }
