/* definitions.h as PPI's CMake configures it for the default build with -DPROTECTED=ON. */
#define ERROR(X,Y) fabs((X)-(Y))
#define PROTECTED 1
