.iversion
.version
.malloc 5
qubit q0
cbit c0
H q0
CNOT q0,q1
T q1
Tdag q2
H q2
S q3
X q4
H q3
CNOT q3,q4
PrepZ q0
MeasZ q0
MeasZ q1
MeasZ q2
MeasZ q3
MeasZ q4
Foo bar
*
.malloc 3
.free
.free
.malloc 100
.malloc abc
