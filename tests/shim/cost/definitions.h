/* definitions.h for a classification build: clamped COST error, REDUCEMAX, PROTECTED. */
#define ERROR(X,Y) COST[(int)fminf(fmaxf((X),0.0f),3.0f)][(int)(Y)]
#define REDUCEMAX 1
#define PROTECTED 1
