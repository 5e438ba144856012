// tetfit_main.cpp - placeholder replaced below
int main() { return 0; }
