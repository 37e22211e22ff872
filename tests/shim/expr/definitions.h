/* definitions.h as PPI's CMake configures it with -DPROTECTED=ON and a user error function,
   -DEF='(((X)-(Y))*((X)-(Y))/(1.0f+fabs(Y)))' (CMakeLists.txt:26-33): a relative squared error. */
#define ERROR(X,Y) (((X)-(Y))*((X)-(Y))/(1.0f+fabs(Y)))
#define PROTECTED 1
