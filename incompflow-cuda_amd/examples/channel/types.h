#ifndef TYPES_H
#define TYPES_H

using scalar = float;

#endif
