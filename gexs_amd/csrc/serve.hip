// serve.hip -- continuous-batching request server (see serve.h).
#include "serve.h"
