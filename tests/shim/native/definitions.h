/* definitions.h as PPI's CMake configures it with -DPROTECTED=ON -DNATIVE=ON (CMakeLists.txt:35-52). */
#define ERROR(X,Y) fabs((X)-(Y))
#define PROTECTED 1
#define NATIVE 1
