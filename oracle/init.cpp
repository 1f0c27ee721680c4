// init.cpp -- placeholder
